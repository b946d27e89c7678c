"""Error behaviour of the C ABI (argument validation happens on the host before
any launch, so every case here returns an error status and leaves the context
usable): bad point data, bad mesh data, bad solver arguments.  Each failure
is reported as a negative
PSSFSS_E* status with a message in pssfss_last_error, which the Python layer
raises as HipError."""
import math
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EINVAL = -1


def _say(msg):
    print(f"\n[case] {msg}", flush=True)


def _strip(pk):
    sh = pk.rectstrip(Nx=4, Ny=4, Lx=5, Ly=5, Px=10, Py=10, units="mm")
    return sh, pk.setup_rwg(sh)


def test_point_validation(ctx, pk):
    _say("errors: point data (nl, s, k0, u, null layer arrays)")
    from pssfss_amd import hip
    sh, rw = _strip(pk)
    ds = hip.DeviceSheet(ctx, sh, rw)
    L = pk.Layer.make
    k0 = 2 * math.pi * 10e9 / 299792458.0
    u = 0.5 * max(np.linalg.norm(sh.beta1), np.linalg.norm(sh.beta2)) * sh.upm
    pb = hip.PointBuilder([L(), L()], sh, u, 1, k0, np.zeros(2), with_modes=False)
    Z, _ = ds.fill(pb)                      # the valid point works
    assert np.isfinite(Z).all()
    nbf = rw.nbf
    zbuf = np.zeros(2 * nbf * nbf)
    cases = [("nl", 1, "nl must be in"), ("s", 0, "s must be in"), ("s", 2, "s must be in"),
             ("k0", 0.0, "k0 and u must be positive"), ("u", -1.0, "k0 and u must be positive")]
    for field, bad, msg in cases:
        old = getattr(pb.point, field)
        setattr(pb.point, field, bad)
        rc = hip.lib().pssfss_fill(ctx.h, ds.h, C.byref(pb.point), hip._d(zbuf), None)
        err = hip.lib().pssfss_last_error(ctx.h).decode()
        setattr(pb.point, field, old)
        print(f"  {field}={bad}: status {rc}, '{err}'")
        assert rc == EINVAL and msg in err
    pb.point.width = None
    rc = hip.lib().pssfss_fill(ctx.h, ds.h, C.byref(pb.point), hip._d(zbuf), None)
    pb.point.width = hip._d(pb.width)
    assert rc == EINVAL and "null layer arrays" in hip.lib().pssfss_last_error(ctx.h).decode()
    with pytest.raises(hip.HipError, match="null argument"):
        ctx.check(hip.lib().pssfss_fill(ctx.h, ds.h, C.byref(pb.point), None, None))
    Z2, _ = ds.fill(pb)                     # the context is still usable
    assert np.array_equal(Z, Z2)
    ds.free()


def test_mesh_validation(ctx, pk):
    _say("errors: sheet upload with inconsistent mesh data")
    from pssfss_amd import hip
    sh, rw = _strip(pk)
    fv = np.array(sh.fv, dtype=np.int64, copy=True)
    fv[0, 0] = sh.rho.shape[0] + 5          # vertex index out of range
    sh.fv = fv
    with pytest.raises(hip.HipError, match="fv index out of range"):
        hip.DeviceSheet(ctx, sh, rw)
    sh2, rw2 = _strip(pk)
    sh2.sheet_class = "X"
    with pytest.raises(hip.HipError, match="class must be"):
        hip.DeviceSheet(ctx, sh2, rw2)


def test_solver_validation(ctx):
    _say("errors: batched solve arguments")
    from pssfss_amd import hip
    A = np.zeros(8)
    B = np.zeros(4)
    it = (C.c_int * 1)()
    rs = np.zeros(1)
    for n, nrhs, nb, prec, msg in [(0, 1, 1, 0, "bad arguments"), (2, 0, 1, 0, "bad arguments"),
                                   (2, 1, 0, 0, "bad arguments"), (2, 1, 1, 7, "precision")]:
        rc = hip.lib().pssfss_lu_solve_batch(ctx.h, n, nrhs, nb, hip._d(A), hip._d(B), prec, it,
                                             hip._d(rs))
        err = hip.lib().pssfss_last_error(ctx.h).decode()
        print(f"  n={n} nrhs={nrhs} batch={nb} precision={prec}: status {rc}, '{err}'")
        assert rc == EINVAL and msg in err
    X, _, _ = ctx.lu_solve(np.eye(3, dtype=complex) * 2, np.ones((3, 1), dtype=complex))
    assert np.allclose(X, 0.5)
