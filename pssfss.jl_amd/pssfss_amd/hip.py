"""ctypes binding of libpssfss_hip.so (include/pssfss_hip.h) and the batched
sheet solver used by `analyze`.

The library is mandatory: importing this module on a machine where the built
library is missing raises, and every entry point raises `HipError` on a
non-zero status with the library's own error text.
"""
import ctypes as C
import math
import os
import time
from typing import List

import numpy as np

from . import gsm as G
from . import outputs as O
from . import modes as Md
from .constants import TWOPI

_HERE = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_HERE, "libpssfss_hip.so")

FP64 = 0
MIXED = 1
PRECISION = {"fp64": FP64, "mixed": MIXED}


class HipError(RuntimeError):
    pass


_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int64)


class SheetDesc(C.Structure):
    _fields_ = [("nnode", C.c_int64), ("nface", C.c_int64), ("nedge", C.c_int64),
                ("nbf", C.c_int64), ("rho", _dp), ("fv", _ip), ("fe", _ip), ("fr", _dp),
                ("bff", _ip), ("ebf", _ip), ("eci", _ip), ("nufp", C.c_int64),
                ("ufp_offsets", _ip), ("ufp_members", _ip), ("s1", C.c_double * 2),
                ("s2", C.c_double * 2), ("beta1", C.c_double * 2), ("beta2", C.c_double * 2),
                ("units_per_meter", C.c_double), ("sheet_class", C.c_char)]


class Point(C.Structure):
    _fields_ = [("k0", C.c_double), ("u", C.c_double), ("psi1", C.c_double),
                ("psi2", C.c_double), ("nl", C.c_int32), ("s", C.c_int32), ("epsr", _dp),
                ("mur", _dp), ("width", _dp), ("nmodes", C.c_int32 * 2), ("beta", _dp * 2),
                ("tvec", _dp * 2), ("c", _dp * 2), ("Y", _dp * 2), ("nmax", C.c_int32),
                ("vinc", _dp), ("tlgf", _dp), ("acf", C.c_double * 2), ("area", C.c_double)]


class Gsm(C.Structure):
    _fields_ = [("s11", _dp), ("s12", _dp), ("s21", _dp), ("s22", _dp)]


class Stats(C.Structure):
    _fields_ = [("ms", C.c_double * 8), ("flops", C.c_double * 8), ("launches", C.c_int64 * 8)]


STAGES = ("spatial_sums", "tables", "fill", "lu_panel", "lu_gemm", "solve", "transforms_rhs",
          "lu_total_flops")


class PointInfo(C.Structure):
    _fields_ = [("mmax", C.c_int32), ("table_converged", C.c_int32),
                ("jk_rings_max", C.c_int32), ("refine_iters", C.c_int32),
                ("fallback_fp64", C.c_int32), ("pad", C.c_int32), ("refine_resid", C.c_double)]


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise HipError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.pssfss_ctx_create.argtypes = [C.c_int, C.POINTER(vp)]
        L.pssfss_ctx_destroy.argtypes = [vp]
        L.pssfss_last_error.argtypes = [vp]
        L.pssfss_last_error.restype = C.c_char_p
        L.pssfss_device_info.argtypes = [vp, C.c_char_p, C.c_int, C.POINTER(C.c_int),
                                         C.POINTER(C.c_double)]
        L.pssfss_sheet_upload.argtypes = [vp, C.POINTER(SheetDesc), C.POINTER(vp)]
        L.pssfss_sheet_free.argtypes = [vp, vp]
        L.pssfss_sheet_gsm.argtypes = [vp, vp, C.POINTER(Point), C.POINTER(Gsm), C.c_int,
                                       C.POINTER(PointInfo)]
        L.pssfss_sheets_gsm_batch.argtypes = [vp, C.c_int, C.POINTER(vp), C.POINTER(C.c_int),
                                              C.POINTER(Point), C.POINTER(Gsm), C.c_int,
                                              C.POINTER(PointInfo)]
        L.pssfss_sheet_gsm_batch.argtypes = [vp, vp, C.c_int, C.POINTER(Point), C.POINTER(Gsm),
                                             C.c_int, C.POINTER(PointInfo)]
        L.pssfss_sheet_prefetch_jk.argtypes = [vp, vp, C.c_int, _dp, _dp, _dp, C.c_int]
        L.pssfss_sheet_adjoint.argtypes = [vp, vp, C.c_int, C.POINTER(Point), C.c_int, _dp,
                                           C.c_int, _dp, _dp, C.c_int]
        L.pssfss_face_pair_classes.argtypes = [vp, C.c_int64, _dp, C.c_int64, _ip,
                                               C.POINTER(C.c_int64), _ip, _ip, _ip]
        L.pssfss_fill.argtypes = [vp, vp, C.POINTER(Point), _dp, C.POINTER(PointInfo)]
        L.pssfss_fill_dk0.argtypes = [vp, vp, C.POINTER(Point), _dp, _dp, C.POINTER(PointInfo)]
        L.pssfss_filljk.argtypes = [vp, vp, C.c_double, C.c_double, C.c_double, _dp, _dp, _dp]
        L.pssfss_modal_tables.argtypes = [vp, C.c_int, C.POINTER(Point), _dp, _dp, _dp,
                                          C.c_double, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                          _dp, _dp]
        L.pssfss_rwgbfft.argtypes = [vp, vp, _dp, C.c_double, C.c_double, _dp]
        L.pssfss_lu_solve_batch.argtypes = [vp, C.c_int64, C.c_int, C.c_int, _dp, _dp, C.c_int,
                                            C.POINTER(C.c_int), _dp]
        L.pssfss_cascade_batch.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp]
        L.pssfss_lu_solve_batch_noalloc.argtypes = [vp, C.c_int64, C.c_int, C.c_int, _dp, _dp]
        L.pssfss_ctx_set_chunk.argtypes = [vp, C.c_int, C.c_double]
        L.pssfss_stats_enable.argtypes = [vp, C.c_int]
        L.pssfss_stats_read.argtypes = [vp, C.POINTER(Stats), C.c_int]
        for name in ("pssfss_ctx_create", "pssfss_ctx_destroy", "pssfss_device_info",
                     "pssfss_sheet_upload", "pssfss_sheet_free", "pssfss_sheet_gsm",
                     "pssfss_sheet_gsm_batch", "pssfss_fill", "pssfss_fill_dk0", "pssfss_filljk",
                     "pssfss_modal_tables", "pssfss_rwgbfft", "pssfss_lu_solve_batch",
                     "pssfss_cascade_batch", "pssfss_stats_enable", "pssfss_stats_read",
                     "pssfss_lu_solve_batch_noalloc", "pssfss_ctx_set_chunk"):
            getattr(L, name).restype = C.c_int
        _LIB = L
    return _LIB


EXPORTS = ("pssfss_ctx_create", "pssfss_ctx_destroy", "pssfss_last_error", "pssfss_device_info",
           "pssfss_sheet_upload", "pssfss_sheet_free", "pssfss_sheet_gsm",
           "pssfss_sheet_gsm_batch", "pssfss_fill", "pssfss_filljk", "pssfss_modal_tables",
           "pssfss_rwgbfft", "pssfss_lu_solve_batch", "pssfss_cascade_batch",
           "pssfss_stats_enable", "pssfss_stats_read", "pssfss_sheet_adjoint",
           "pssfss_face_pair_classes", "pssfss_fill_dk0", "pssfss_lu_solve_batch_noalloc",
           "pssfss_ctx_set_chunk", "pssfss_sheets_gsm_batch", "pssfss_sheet_prefetch_jk")


def _d(a):
    return a.ctypes.data_as(_dp)


def _cplx(a) -> np.ndarray:
    """Private contiguous float64 (re, im) copy of a complex array."""
    return np.array(a, dtype=np.complex128, order="C", copy=True).view(np.float64)


class Context:
    def __init__(self, device: int = 0):
        self.h = C.c_void_p()
        st = lib().pssfss_ctx_create(device, C.byref(self.h))
        if st != 0:
            raise HipError(f"pssfss_ctx_create failed with status {st}")

    def check(self, st):
        if st != 0:
            msg = lib().pssfss_last_error(self.h).decode()
            raise HipError(f"status {st}: {msg}")

    def device_info(self):
        name = C.create_string_buffer(256)
        cu = C.c_int()
        gib = C.c_double()
        self.check(lib().pssfss_device_info(self.h, name, 256, C.byref(cu), C.byref(gib)))
        return name.value.decode(), cu.value, gib.value

    def set_chunk(self, max_points: int = 0, mem_fraction: float = 0.0):
        """Points per chunk of the batched sheet calls and their share of free HBM
        (0: the defaults, 128 points and 0.45)."""
        self.check(lib().pssfss_ctx_set_chunk(self.h, int(max_points), float(mem_fraction)))

    def stats_enable(self, on: bool = True):
        self.check(lib().pssfss_stats_enable(self.h, 1 if on else 0))

    def stats_read(self, reset: bool = True) -> dict:
        st = Stats()
        self.check(lib().pssfss_stats_read(self.h, C.byref(st), 1 if reset else 0))
        return {name: {"ms": st.ms[i], "flops": st.flops[i], "launches": st.launches[i]}
                for i, name in enumerate(STAGES)}

    def close(self):
        if self.h:
            lib().pssfss_ctx_destroy(self.h)
            self.h = C.c_void_p()

    # ---- single-stage entries (parity boundaries)
    def lu_solve(self, A: np.ndarray, B: np.ndarray, precision="fp64"):
        """Solve A[b] X = B[b] for a batch (A: (nb, n, n), B: (nb, n, nrhs))."""
        A = np.asarray(A, np.complex128)
        B = np.asarray(B, np.complex128)
        if A.ndim == 2:
            A, B = A[None], B[None]
        nb, n, _ = A.shape
        nrhs = B.shape[2] if B.ndim == 3 else 1
        # private column-major copies (the library overwrites B with X)
        Af = np.array(np.transpose(A, (0, 2, 1)), order="C", copy=True).view(np.float64)
        Bf = np.array(np.transpose(B.reshape(nb, n, nrhs), (0, 2, 1)), order="C",
                      copy=True).view(np.float64)
        it = (C.c_int * nb)()
        rs = np.zeros(nb)
        self.check(lib().pssfss_lu_solve_batch(self.h, n, nrhs, nb, _d(Af), _d(Bf),
                                               PRECISION[precision], it, _d(rs)))
        X = np.transpose(Bf.view(np.complex128).reshape(nb, nrhs, n), (0, 2, 1))
        return X, list(it), rs

    def cascade_batch(self, a_list, b_list):
        """Device Redheffer star products of GSM pairs (lists of GSM)."""
        nb = len(a_list)
        n1, n2 = a_list[0].s12.shape
        n3 = b_list[0].s22.shape[0]

        def pack(g):
            return np.concatenate([np.asfortranarray(m).ravel(order="F")
                                   for m in (g.s11, g.s12, g.s21, g.s22)])
        A = _cplx(np.concatenate([pack(g) for g in a_list]))
        B = _cplx(np.concatenate([pack(g) for g in b_list]))
        lc = n1 * n1 + 2 * n1 * n3 + n3 * n3
        Cc = np.zeros(2 * lc * nb)
        self.check(lib().pssfss_cascade_batch(self.h, nb, n1, n2, n3, _d(A), _d(B), _d(Cc)))
        out = []
        cz = Cc.view(np.complex128)
        for e in range(nb):
            v = cz[e * lc:(e + 1) * lc]
            o = 0
            mats = []
            for (r, c) in ((n1, n1), (n1, n3), (n3, n1), (n3, n3)):
                mats.append(v[o:o + r * c].reshape((r, c), order="F").copy())
                o += r * c
            out.append(G.GSM(*mats))
        return out


_CTX = None


def get_context() -> Context:
    global _CTX
    if _CTX is None:
        _CTX = Context(0)
    return _CTX


def face_pair_classes(ctx: "Context", sheet):
    """Unique-face-pair classes of `sheet` on the device: (offsets, members, ufpm) as
    rwg.unique_face_pairs returns them (ufpm[m, n] 0-based)."""
    nface = sheet.nface
    npair = nface * nface
    rho = np.ascontiguousarray(sheet.rho, np.float64)
    fv = np.ascontiguousarray((sheet.fv + 1).T, np.int64)        # 3 per face, 1-based
    offsets = np.zeros(npair + 1, np.int64)
    members = np.zeros(npair, np.int64)
    ufpm1 = np.zeros(npair, np.int64)
    n = C.c_int64()
    ctx.check(lib().pssfss_face_pair_classes(ctx.h, rho.shape[0], _d(rho), nface,
                                             fv.ctypes.data_as(_ip), C.byref(n),
                                             offsets.ctypes.data_as(_ip),
                                             members.ctypes.data_as(_ip),
                                             ufpm1.ctypes.data_as(_ip)))
    ufpm = (ufpm1 - 1).reshape(nface, nface).T.copy()
    return offsets[:n.value + 1].copy(), members, ufpm


class DeviceSheet:
    """An RWGSheet + RWGData resident on the device (pssfss_sheet_upload)."""

    def __init__(self, ctx: Context, sheet, rwgdat):
        self.ctx = ctx
        self.sheet = sheet
        self.rwgdat = rwgdat
        self._keep = []
        k = self._keep
        rho = np.ascontiguousarray(sheet.rho, np.float64)
        fv = np.ascontiguousarray(sheet.fv.T + 1, np.int64)
        fe = np.ascontiguousarray(sheet.fe.T + 1, np.int64)
        fr = np.ascontiguousarray(sheet.fr if len(sheet.fr) else np.zeros(sheet.nface), np.float64)
        bff = np.ascontiguousarray(rwgdat.bff.T + 1, np.int64)
        ebf = np.ascontiguousarray(rwgdat.ebf + 1, np.int64)
        eci = np.ascontiguousarray(rwgdat.eci, np.int64)
        k += [rho, fv, fe, fr, bff, ebf, eci]
        d = SheetDesc()
        d.nnode, d.nface, d.nedge, d.nbf = len(rho), sheet.nface, sheet.nedge, rwgdat.nbf
        d.rho, d.fr = _d(rho), _d(fr)
        d.fv, d.fe = fv.ctypes.data_as(_ip), fe.ctypes.data_as(_ip)
        d.bff, d.ebf, d.eci = (bff.ctypes.data_as(_ip), ebf.ctypes.data_as(_ip),
                               eci.ctypes.data_as(_ip))
        d.nufp = rwgdat.nufp
        if rwgdat.ufp_identity:
            d.ufp_offsets = C.cast(None, _ip)
            d.ufp_members = C.cast(None, _ip)
        else:
            off = np.ascontiguousarray(rwgdat.ufp_offsets, np.int64)
            mem = np.ascontiguousarray(rwgdat.ufp_members + 1, np.int64)
            k += [off, mem]
            d.ufp_offsets, d.ufp_members = off.ctypes.data_as(_ip), mem.ctypes.data_as(_ip)
        for i in range(2):
            d.s1[i], d.s2[i] = sheet.s1[i], sheet.s2[i]
            d.beta1[i], d.beta2[i] = sheet.beta1[i], sheet.beta2[i]
        d.units_per_meter = sheet.upm
        d.sheet_class = sheet.sheet_class.encode()
        self.h = C.c_void_p()
        ctx.check(lib().pssfss_sheet_upload(ctx.h, C.byref(d), C.byref(self.h)))
        self._keep = []

    def free(self):
        if self.h:
            self.ctx.check(lib().pssfss_sheet_free(self.ctx.h, self.h))
            self.h = C.c_void_p()

    def filljk(self, u_local, psi1, psi2):
        nu = self.rwgdat.nufp
        jk = np.zeros(12 * nu)
        rr = np.zeros(2 * nu)
        ri = np.zeros(nu)
        self.ctx.check(lib().pssfss_filljk(self.ctx.h, self.h, u_local, psi1, psi2, _d(jk),
                                           _d(rr), _d(ri)))
        return jk.view(np.complex128).reshape(nu, 6), rr.reshape(nu, 2), ri

    def rwgbfft(self, k, psi1, psi2):
        nbf = self.rwgdat.nbf
        ft = np.zeros(4 * nbf)
        kk = np.ascontiguousarray(k, np.float64)
        self.ctx.check(lib().pssfss_rwgbfft(self.ctx.h, self.h, _d(kk), psi1, psi2, _d(ft)))
        return ft.view(np.complex128).reshape(nbf, 2)

    def fill(self, pb: "PointBuilder"):
        nbf = self.rwgdat.nbf
        Z = np.zeros(2 * nbf * nbf)
        info = PointInfo()
        self.ctx.check(lib().pssfss_fill(self.ctx.h, self.h, C.byref(pb.point), _d(Z),
                                         C.byref(info)))
        return Z.view(np.complex128).reshape((nbf, nbf), order="F"), info

    def fill_dk0(self, pb: "PointBuilder"):
        """(Z, dZ/dk0, info) at the point's phasing and u; k0 in 1/m."""
        nbf = self.rwgdat.nbf
        Z = np.zeros(2 * nbf * nbf)
        dZ = np.zeros(2 * nbf * nbf)
        info = PointInfo()
        self.ctx.check(lib().pssfss_fill_dk0(self.ctx.h, self.h, C.byref(pb.point), _d(Z),
                                             _d(dZ), C.byref(info)))
        f = lambda a: a.view(np.complex128).reshape((nbf, nbf), order="F")
        return f(Z), f(dZ), info


class PointBuilder:
    """Holds the arrays of one `pssfss_point` alive and builds the struct."""

    def __init__(self, region, sheet, u, s, k0, kinc, with_modes=True):
        upm = sheet.upm
        self.psi1 = float(kinc @ sheet.s1) / upm
        self.psi2 = float(kinc @ sheet.s2) / upm
        nl = len(region)
        self.eps = _cplx([l.epsr for l in region])
        self.mu = _cplx([l.mur for l in region])
        self.width = np.ascontiguousarray([l.width for l in region], np.float64)
        p = Point()
        p.k0, p.u, p.psi1, p.psi2 = k0, u, self.psi1, self.psi2
        p.nl, p.s = nl, s
        p.epsr, p.mur, p.width = _d(self.eps), _d(self.mu), _d(self.width)
        self.gsm = None
        if with_modes:
            ends = (region[0], region[-1])
            area = abs(sheet.s1[0] * sheet.s2[1] - sheet.s1[1] * sheet.s2[0]) / upm ** 2
            acf = []
            for l in ends:
                ai = TWOPI * TWOPI / abs(l.beta1[0] * l.beta2[1] - l.beta1[1] * l.beta2[0])
                acf.append(math.sqrt(ai / area))
            if sheet.sheet_class == "M":
                g, tl, vi = G.gsm_magnetic_gblock(region, s, k0)
            else:
                g, tl, vi = G.gsm_electric_gblock(region, s, k0)
            self.gsm = G.GSM(*[np.asfortranarray(m) for m in (g.s11, g.s12, g.s21, g.s22)])
            self.arrs = []
            for r, l in enumerate(ends):
                nm = len(l.P)
                p.nmodes[r] = nm
                # private copies: setup_modes overwrites the layer arrays per point
                b = np.array(l.beta[:nm], dtype=np.float64, order="C", copy=True).reshape(-1)
                t = np.array(l.tvec[:nm], dtype=np.float64, order="C", copy=True).reshape(-1)
                c = _cplx(l.c[:nm])
                y = _cplx(l.Y[:nm])
                self.arrs += [b, t, c, y]
                p.beta[r], p.tvec[r], p.c[r], p.Y[r] = _d(b), _d(t), _d(c), _d(y)
            nmax = vi.shape[0]
            self.vinc = _cplx(vi.T)
            self.tlgf = _cplx(tl.T)
            p.nmax = nmax
            p.vinc, p.tlgf = _d(self.vinc), _d(self.tlgf)
            p.acf[0], p.acf[1] = acf
            p.area = area
            self.gsm_arrs = [np.ascontiguousarray(m.ravel(order="F")).view(np.float64)
                             if m.size else np.zeros(2)
                             for m in (self.gsm.s11, self.gsm.s12, self.gsm.s21, self.gsm.s22)]
        self.point = p

    def gsm_struct(self) -> Gsm:
        g = Gsm()
        g.s11, g.s12, g.s21, g.s22 = [_d(a) for a in self.gsm_arrs]
        return g

    def result(self) -> "G.GSM":
        mats = [a.view(np.complex128).reshape(m.shape, order="F").copy()
                for a, m in zip(self.gsm_arrs, (self.gsm.s11, self.gsm.s12,
                                                self.gsm.s21, self.gsm.s22))]
        return G.GSM(*mats)


def sheet_gsm_batch(ctx: Context, dsheet: DeviceSheet, builders: List[PointBuilder],
                    precision: str = "fp64"):
    n = len(builders)
    pts = (Point * n)(*[b.point for b in builders])
    gs = (Gsm * n)(*[b.gsm_struct() for b in builders])
    infos = (PointInfo * n)()
    ctx.check(lib().pssfss_sheet_gsm_batch(ctx.h, dsheet.h, n, pts, gs, PRECISION[precision],
                                           infos))
    return [b.result() for b in builders], list(infos)


def sheets_gsm_batch(ctx: Context, groups, precision: str = "fp64"):
    """Sheet GSMs of several sheets in shared batches (pssfss_sheets_gsm_batch):
    groups is a list of (DeviceSheet, [PointBuilder, ...]); every point must have
    the same end-region mode counts.  Returns per group (gsms, infos)."""
    builders = [b for _, bs in groups for b in bs]
    n = len(builders)
    if n == 0:
        return [([], []) for _ in groups]
    pts = (Point * n)(*[b.point for b in builders])
    gs = (Gsm * n)(*[b.gsm_struct() for b in builders])
    infos = (PointInfo * n)()
    handles = (C.c_void_p * len(groups))(*[d.h for d, _ in groups])
    counts = (C.c_int * len(groups))(*[len(bs) for _, bs in groups])
    ctx.check(lib().pssfss_sheets_gsm_batch(ctx.h, len(groups), handles, counts, pts, gs,
                                            PRECISION[precision], infos))
    out, k = [], 0
    for _, bs in groups:
        out.append(([b.result() for b in bs], list(infos[k:k + len(bs)])))
        k += len(bs)
    return out


class SheetSolver:
    """Runs the `_analyze` sweep with every sheet GSM computed on the device,
    all points of one sheet Gblock in one batched call."""

    # largest GSM port the device cascade kernel takes
    CASCADE_MAX_PORT = 34

    def __init__(self, ctx: Context, setup, precision: str = "fp64", max_batch: int = 0,
                 cascade: str = "device"):
        self.ctx = ctx
        self.setup = setup
        self.precision = precision
        self.max_batch = max_batch      # points per library call; 0 = all points of a sheet
        self.cascade = cascade          # "device": batched star products on the GPU; "host"
        self.dsheets = {}

    def upload(self):
        """Upload every analysed sheet now (otherwise the first run() does it)."""
        for i, sh in enumerate(self.setup.sheets):
            if sh.style != "NULL" and sh.sheet_class in ("J", "M"):
                self._dsheet(i)

    def _dsheet(self, isheet):
        if isheet not in self.dsheets:
            su = self.setup
            self.dsheets[isheet] = DeviceSheet(self.ctx, su.sheets[isheet], su.rwgdat[isheet])
        return self.dsheets[isheet]

    def run(self, limit: int = 0):
        """The sweep (its first `limit` analysed points when limit > 0)."""
        self.prefetch(limit)
        points, live, requests = self.requests(limit)
        solved = {}
        infos_all = []
        for ig, reqs in requests.items():
            ds = self.sheet_of(ig)
            step = self.max_batch if self.max_batch > 0 else len(reqs)
            for r0 in range(0, len(reqs), step):
                part = reqs[r0:r0 + step]
                gsms, infos = sheet_gsm_batch(self.ctx, ds, [b for _, b in part], self.precision)
                for (ip, _), g, inf in zip(part, gsms, infos):
                    solved[(ip, ig)] = g
                infos_all += infos
        return self.finish(points, live, solved, infos_all)

    def prefetch(self, limit: int = 0):
        """Start every analysed sheet's spatial sums before the points are prepared on
        the host (pssfss_sheet_prefetch_jk), with the phasings and u of exactly the
        points run() will pass (the cut-off skip applied, as PointBuilder sets them).
        One key: into the sheet's cache; several: the GSM call's plan, used when all
        of a sheet's points go to the library in one call."""
        su = self.setup
        live = []
        for pt in su.points():
            if limit > 0 and len(live) >= limit:
                break
            for l in su.layers:
                Md.setup_modes(l, pt.k0, pt.beta00)
            if su.principal_ok():
                live.append(pt)
        if not live:
            return
        if self.max_batch > 0 and self.max_batch < len(live):
            return
        # sheets in the order run() solves them; a plan of several keys occupies the
        # context's one set of slots, so only the first sheet's is started
        order = []
        for ig, gbl in enumerate(su.gbls):
            if gbl.j == 0 or su.gbldup[ig] > 0:
                continue
            ish = su.junc[gbl.j - 1]
            sheet = su.sheets[ish - 1]
            if sheet.style != "NULL" and sheet.sheet_class in ("J", "M") and ish not in order:
                order.append(ish)
        for k, ish in enumerate(order):
            sheet = su.sheets[ish - 1]
            psi1 = np.array([float(p.beta00 @ sheet.s1) / sheet.upm for p in live])
            psi2 = np.array([float(p.beta00 @ sheet.s2) / sheet.upm for p in live])
            if k > 0 and len(set(zip(psi1.tolist(), psi2.tolist()))) > 1:
                continue
            ds = self._dsheet(ish - 1)
            u = np.full(len(live), float(su.uvec[ish - 1]))
            self.ctx.check(lib().pssfss_sheet_prefetch_jk(self.ctx.h, ds.h, len(live),
                                                          _d(psi1), _d(psi2), _d(u),
                                                          PRECISION[self.precision]))

    def sheet_of(self, ig):
        """Device sheet of sheet Gblock ig."""
        su = self.setup
        return self._dsheet(su.junc[su.gbls[ig].j - 1] - 1)

    def requests(self, limit: int = 0):
        """The sweep's analysed points (the cut-off skip of src/PSSFSS.jl:252-255) and,
        per sheet Gblock, the list of (point index, PointBuilder) to solve."""
        su = self.setup
        points = list(su.points())
        live = []
        requests = {}    # (gbl index) -> list of (point index, builder)
        for ip, pt in enumerate(points):
            if limit > 0 and len(live) >= limit:
                break
            for l in su.layers:
                Md.setup_modes(l, pt.k0, pt.beta00)
            if not su.principal_ok():
                continue
            live.append(ip)
            for ig, gbl in enumerate(su.gbls):
                if gbl.j == 0 or su.gbldup[ig] > 0:
                    continue
                ish = su.junc[gbl.j - 1]
                sheet = su.sheets[ish - 1]
                if sheet.sheet_class not in ("J", "M"):
                    continue
                i1, i2 = gbl.rng[0], 1 + gbl.rng[-1]
                region = su.layers[i1 - 1:i2]
                s = gbl.j - i1 + 1
                if sheet.style == "NULL":
                    continue
                b = PointBuilder(region, sheet, su.uvec[ish - 1], s, pt.k0, pt.beta00)
                requests.setdefault(ig, []).append((ip, b))
        return points, live, requests

    def finish(self, points, live, solved, infos_all):
        """Cascade every analysed point from its solved sheet GSMs (keys (point,
        Gblock)) and return (results, stats) as run() does."""
        su = self.setup
        t0 = time.perf_counter()
        ports = [len(l.P) for l in su.layers]
        if self.cascade == "device" and live and max(ports) <= self.CASCADE_MAX_PORT:
            totals = self._cascade_device(points, live, solved)
        else:
            totals = self._cascade_host(points, live, solved)
        results = [O.make_result(g, points[ip].steer, points[ip].beta00, points[ip].fghz,
                                 su.layers) for ip, g in zip(live, totals)]
        stats = {"npoints": len(live), "infos": infos_all,
                 "cascade_s": time.perf_counter() - t0}
        return results, stats

    def _cascade_host(self, points, live, solved):
        """The reference's walk (src/PSSFSS.jl:256-305) point by point on the host."""
        su = self.setup
        out = []
        for ip in live:
            pt = points[ip]
            for l in su.layers:
                Md.setup_modes(l, pt.k0, pt.beta00)
            gbl_of = {}

            def solver(region, sheet, u, rwgdat, s, isheet, _ip=ip):
                for ig, gbl in enumerate(su.gbls):
                    if gbl.j != 0 and su.junc[gbl.j - 1] - 1 == isheet and (_ip, ig) in solved \
                            and (_ip, ig) not in gbl_of:
                        gbl_of[(_ip, ig)] = True
                        return solved[(_ip, ig)].copy()
                raise RuntimeError("sheet GSM missing from the batched solve")
            out.append(su.cascade_point(pt.k0, solver))
        return out

    def _cascade_device(self, points, live, solved):
        """The same walk with every point advanced together: per Gblock, one batched
        star product (src/GSMs.jl:125-141) of all points on the device
        (pssfss_cascade_batch), then the propagation through the next layer
        (src/GSMs.jl:155-176) as diagonal scalings."""
        su = self.setup
        layers = su.layers
        nb = len(live)
        blocks = [[None] * nb for _ in su.gbls]
        gammas = []
        for k, ip in enumerate(live):
            pt = points[ip]
            for l in layers:
                Md.setup_modes(l, pt.k0, pt.beta00)
            gammas.append([l.gamma.copy() for l in layers])
            saved = {}
            for ig, gbl in enumerate(su.gbls):
                i1, i2 = gbl.rng[0], 1 + gbl.rng[-1]
                if gbl.j == 0:
                    blocks[ig][k] = G.gsm_slab_interface(layers[i1 - 1], layers[i2 - 1], pt.k0)
                    continue
                sheet = su.sheets[su.junc[gbl.j - 1] - 1]
                region = layers[i1 - 1:i2]
                if su.gbldup[ig] > 0:
                    gb = saved[su.gbldup[ig] - 1].copy()
                else:
                    if sheet.sheet_class in ("J", "M"):
                        gb = solved[(ip, ig)].copy()
                    elif sheet.sheet_class == "E":
                        gb = G.pecgsm(len(region[0].P), len(region[-1].P))
                    elif sheet.sheet_class == "H":
                        gb = G.pmcgsm(len(region[0].P), len(region[-1].P))
                    else:
                        raise ValueError(f"Illegal sheet class: {sheet.sheet_class}")
                    if su.gbldup[ig] < 0:
                        saved[ig] = gb.copy()
                if sheet.dx != 0 or sheet.dy != 0:
                    G.translate_gsm(gb, sheet.dx / sheet.upm, sheet.dy / sheet.upm, region[0],
                                    region[-1])
                blocks[ig][k] = gb

        def propagate(gs, il):
            w = layers[il].width
            if w == 0:
                return
            for k, g in enumerate(gs):
                p = np.exp(-gammas[k][il] * w)
                g.s12 *= p[None, :]
                g.s21 *= p[:, None]
                g.s22 *= p[None, :]
                g.s22 *= p[:, None]
        n1 = len(layers[0].P)
        cur = [G.GSM.new(n1, n1) for _ in live]
        propagate(cur, 0)
        for ig, gbl in enumerate(su.gbls):
            cur = self.ctx.cascade_batch(cur, blocks[ig])
            propagate(cur, gbl.rng[-1])
        return cur

    def close(self):
        for d in self.dsheets.values():
            d.free()
        self.dsheets = {}

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
        return False
