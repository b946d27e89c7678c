"""Device-state lifetimes through the C ABI: the sheet's J/K cache across
precisions, and device memory across repeated analyses and context teardown."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"\n[case] {msg}", flush=True)


def _max_dS(res, ref):
    m = 0.0
    for a, b in zip(res, ref):
        for blk in ("s11", "s12", "s21", "s22"):
            m = max(m, np.abs(getattr(a["gsm"], blk) - getattr(b["gsm"], blk)).max())
    return m


def _free_bytes():
    hip = C.CDLL("libamdhip64.so")
    free, total = C.c_size_t(), C.c_size_t()
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value, total.value


@pytest.mark.parametrize("order", [("mixed", "fp64"), ("fp64", "mixed")])
def test_jk_cache_precision(ctx, pk, orc, order):
    """One device sheet at normal incidence (a single (psi1, psi2, u) key) solved in
    both precisions one after the other: the FP64 sweep must not reuse spatial
    sums summed in FP32 by the mixed sweep, and matches the oracle to 1e-9."""
    _say(f"J/K cache across precisions on one device sheet: {order[0]} then {order[1]}")
    from pssfss_amd.hip import SheetSolver
    strata, freqs, steer = pk.workloads.patch_free_space(n=8, nfreq=3)
    ref = orc.analyze(pk, strata, freqs, steer)
    setup = pk.analyze.prepare(strata, freqs, steer)
    with SheetSolver(ctx, setup, order[0]) as solver:
        first, _ = solver.run()
        assert len(solver.dsheets) == 1
        solver.precision = order[1]
        second, _ = solver.run()
    by = {order[0]: first, order[1]: second}
    d64 = _max_dS(by["fp64"], ref)
    dmx = _max_dS(by["mixed"], ref)
    print(f"  max |dS| fp64 {d64:.3e}, mixed {dmx:.3e}")
    assert d64 <= 1e-9
    assert dmx <= 1e-6


def test_analyze_releases_device_memory(ctx, pk):
    """Twenty consecutive analyses leave the device's free memory where the first
    left it (within 1 %): every call frees its device sheets."""
    _say("device memory across 20 analyze calls")
    strata, freqs, steer = pk.workloads.patch_free_space(n=10, nfreq=2)
    pk.analyze.analyze(strata, freqs, steer, precision="mixed", ctx=ctx)
    f0, total = _free_bytes()
    for _ in range(20):
        pk.analyze.analyze(strata, freqs, steer, precision="mixed", ctx=ctx)
    f1, _ = _free_bytes()
    print(f"  free before {f0 / 2**30:.3f} GiB, after {f1 / 2**30:.3f} GiB of {total / 2**30:.0f}")
    assert abs(f1 - f0) <= 0.01 * f0


def test_ctx_destroy_frees_sheets(pk):
    """Sheets still uploaded when their context is destroyed are freed with it, and
    freeing a sheet twice is an argument error, not a double free."""
    _say("context teardown with live sheets; double free rejected")
    from pssfss_amd import hip
    sh = pk.rectstrip(Lx=5, Ly=5, Px=10, Py=10, Nx=12, Ny=12, units="mm")
    rw = pk.setup_rwg(sh)
    c = hip.Context(0)
    try:
        d = hip.DeviceSheet(c, sh, rw)
        h = C.c_void_p(d.h.value)
        d.free()
        st = hip.lib().pssfss_sheet_free(c.h, h)
        assert st == -1, st
        assert "not live" in hip.lib().pssfss_last_error(c.h).decode()
    finally:
        c.close()
    f0, _ = _free_bytes()
    for _ in range(5):
        c = hip.Context(0)
        for _ in range(4):
            hip.DeviceSheet(c, sh, rw).filljk(0.3, 0.0, 0.0)   # cache filled, never freed
        c.close()
    f1, _ = _free_bytes()
    print(f"  free before {f0 / 2**30:.3f} GiB, after {f1 / 2**30:.3f} GiB")
    assert abs(f1 - f0) <= 0.01 * f0


@pytest.mark.parametrize("steer", [{"theta": [0.0], "phi": [0.0]},
                                   {"theta": [20.0], "phi": [0.0]}])
def test_prefetch_jk_is_transparent(ctx, pk, steer):
    """pssfss_sheet_prefetch_jk starts a sweep's spatial sums before the points are
    prepared; the sweep's S must be bit-identical to a sweep that sums them inside
    the GSM call (normal incidence: one key, into the sheet's cache; oblique: one
    key per frequency, the GSM call's plan in the context's slots)."""
    from pssfss_amd import hip
    _say(f"prefetched spatial sums vs summed in the GSM call, {steer}")
    strata, freqs, _ = pk.workloads.patch_free_space(n=8, nfreq=4)
    su = pk.analyze.prepare(strata, freqs, steer)
    for prec in ("fp64", "mixed"):
        ctx.stats_read(reset=True)
        ctx.stats_enable(True)
        try:
            with hip.SheetSolver(ctx, su, prec) as s:
                res_a, _ = s.run()
        finally:
            ctx.stats_enable(False)
        # one spatial-sum launch (the prefetch's): the GSM call used its integrals
        launches = ctx.stats_read(reset=True)["spatial_sums"]["launches"]
        with hip.SheetSolver(ctx, su, prec) as s:
            s.prefetch = lambda *a, **k: None
            res_b, _ = s.run()
        d = _max_dS(res_a, res_b)
        print(f"  {prec}: max |dS| prefetched vs not = {d:.3e}; spatial-sum launches {launches}")
        assert len(res_a) == len(res_b) == 4
        assert d == 0.0
        assert launches == 1
    ds = hip.DeviceSheet(ctx, su.sheets[0], su.rwgdat[0])
    try:
        one = np.ones(1)
        with pytest.raises(hip.HipError, match="u must be positive"):
            ctx.check(hip.lib().pssfss_sheet_prefetch_jk(ctx.h, ds.h, 1, hip._d(one),
                                                         hip._d(one), hip._d(-one),
                                                         hip.PRECISION["fp64"]))
    finally:
        ds.free()
