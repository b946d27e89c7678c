"""Adjoint geometry gradients (no reference counterpart) against central finite
differences: of the GPU FP64 path, and of the oracle (CPU restatement) for two
of the parameters.  Objective |S21|^2 co-polar TE and TM of a corner-parameterised
patch on the two-layer stack (eps_r 3.0, 1.0 mm; eps_r 1.1 foam, 3.0 mm) at normal
incidence.

The reference algorithm is piecewise smooth: its 6-point table interpolant jumps
across table cells and the spatial sums stop at a 1e-8 ring test, so S moves by
~1e-9 when a quadrature point crosses a cell.  Central differences with step h
then carry an error ~1e-9/h (about 2e-3 relative at h = 1e-6 of the period,
falling as 1/h to 2.5e-5 at 1e-4).  Larger steps cross proportionally more
cells, which leaves a floor near 1e-5.  The check uses the Richardson-extrapolated
central difference (4 D(h) - D(2h)) / 3 at h = 1e-3 of the period (O(h^4)
truncation): max relative error <= 3e-5 over all points, objectives and
parameters, median <= 5e-6."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CORNERS = np.array([[2.1, 2.4], [9.7, 2.0], [10.1, 9.6], [2.3, 9.9]])
P = 12.0
OBJ = [(2, 1, 1, 1), (2, 1, 2, 2)]


def _stack(pk, corners, n=5):
    from pssfss_amd.adjoint import bilinear_patch
    sh, vt = bilinear_patch(corners, n, n, P, P)
    L = pk.Layer.make
    strata = [L(), sh, L(width=1.0, units="mm", epsr=3.0), L(width=3.0, units="mm", epsr=1.1),
              L()]
    return strata, vt


def _objectives(results):
    out = []
    for r in results:
        g = r["gsm"]
        blk = {(1, 1): g.s11, (1, 2): g.s12, (2, 1): g.s21, (2, 2): g.s22}
        out.append([abs(blk[(i, j)][m - 1, n - 1]) ** 2 for (i, j, m, n) in OBJ])
    return np.array(out)


def _cd(pk, analyze, k, freqs, steer, h):
    c1 = CORNERS.copy().reshape(-1)
    c2 = CORNERS.copy().reshape(-1)
    c1[k] += h
    c2[k] -= h
    s1, _ = _stack(pk, c1.reshape(4, 2))
    s2, _ = _stack(pk, c2.reshape(4, 2))
    return (_objectives(analyze(s1, freqs, steer)) - _objectives(analyze(s2, freqs, steer))) / (2 * h)


def _fd(pk, analyze, k, freqs, steer, h):
    return (4 * _cd(pk, analyze, k, freqs, steer, h) - _cd(pk, analyze, k, freqs, steer, 2 * h)) / 3


def test_adjoint_vs_finite_differences(ctx, pk):
    from pssfss_amd.adjoint import gradients
    freqs = [8.0, 12.0]
    steer = {"theta": [0.0], "phi": [0.0]}
    strata, vt = _stack(pk, CORNERS)
    print("\n[case] adjoint gradients, 5x5 bilinear patch, 8 parameters, 2 freqs", flush=True)
    grad, vals = gradients(strata, freqs, steer, vt, OBJ, ctx=ctx)
    assert grad.shape == (2, 2, 8)
    fwd = _objectives(pk.analyze.analyze(strata, freqs, steer, precision="fp64", ctx=ctx))
    assert np.abs(vals - fwd).max() < 1e-12
    h = 1e-3 * P

    def gpu(s, f, st):
        return pk.analyze.analyze(s, f, st, precision="fp64", ctx=ctx)
    fd = np.stack([_fd(pk, gpu, k, freqs, steer, h) for k in range(8)], axis=-1)
    scale = np.abs(fd).max(axis=-1, keepdims=True)
    rel = np.abs(grad - fd) / scale
    print(f"  max |grad - FD| / max|FD| = {rel.max():.2e}, median {np.median(rel):.2e}; "
          f"max|FD| = {scale.max():.3e}")
    assert rel.max() < 3e-5
    assert np.median(rel) < 5e-6


def test_adjoint_vs_oracle_finite_differences(ctx, pk, orc):
    from pssfss_amd.adjoint import gradients
    freqs = [10.0]
    steer = {"theta": [0.0], "phi": [0.0]}
    strata, vt = _stack(pk, CORNERS)
    grad, _ = gradients(strata, freqs, steer, vt, OBJ, ctx=ctx)
    h = 1e-3 * P

    def cpu(s, f, st):
        return orc.analyze(pk, s, f, st)
    for k in (0, 5):
        fd = _fd(pk, cpu, k, freqs, steer, h)
        rel = np.abs(grad[:, :, k] - fd) / np.abs(fd).max()
        print(f"  parameter {k}: |grad - FD(oracle)| / |FD| = {rel.max():.2e}")
        assert rel.max() < 3e-5


def test_adjoint_mixed_vs_fp64(ctx, pk):
    """The mixed form (FP32 LU, FP64 refinement of the forward and the transposed
    systems) against the FP64 adjoint on the same points: gradients agree to 1e-6
    relative to the largest gradient of each objective and point."""
    from pssfss_amd.adjoint import gradients
    freqs = [8.0, 12.0]
    steer = {"theta": [0.0], "phi": [0.0]}
    strata, vt = _stack(pk, CORNERS)
    print("\n[case] mixed vs fp64 adjoint, 5x5 bilinear patch, 8 parameters, 2 freqs", flush=True)
    g64, v64 = gradients(strata, freqs, steer, vt, OBJ, ctx=ctx)
    g32, v32 = gradients(strata, freqs, steer, vt, OBJ, ctx=ctx, precision="mixed")
    rel = np.abs(g32 - g64) / np.abs(g64).max(axis=-1, keepdims=True)
    print(f"  max |grad_mixed - grad_fp64| / max|grad| = {rel.max():.2e}; "
          f"max |f_mixed - f_fp64| = {np.abs(v32 - v64).max():.2e}")
    assert rel.max() < 1e-6
    assert np.abs(v32 - v64).max() < 1e-6
