"""Output functions (src/Outputs.jl) against physics that fixes them independently:
circular-polarisation handedness on transmission and on PEC reflection, axial
ratios from the polarisation-ellipse semi-axes, and steering angles recovered
from the incident wavevector."""
import math

import numpy as np
import pytest


def _result(pk, s11, s21, theta=0.0, phi=0.0, fghz=10.0):
    from pssfss_amd.outputs import Result
    k0 = 2 * math.pi * fghz * 1e9 / 299792458.0
    b00 = k0 * math.sin(math.radians(theta)) * np.array([math.cos(math.radians(phi)),
                                                         math.sin(math.radians(phi))])
    g = pk.GSM(np.asarray(s11, complex), np.asarray(s21, complex).T.copy(),
               np.asarray(s21, complex), np.asarray(s11, complex))
    return Result(g, {"theta": theta, "phi": phi}, b00, fghz)


def _ellipse_ar_db(e):
    """Axial ratio of a 2-component complex field from the semi-axes of the
    ellipse traced by Re(e exp(j w t)): singular values of [Re e, Im e]."""
    m = np.array([[e[0].real, e[0].imag], [e[1].real, e[1].imag]])
    sv = np.linalg.svd(m, compute_uv=False)
    return math.inf if sv[1] == 0 else 20 * math.log10(sv[0] / sv[1])


def test_identity_keeps_handedness(pk):
    from pssfss_amd.outputs import getsijmn, R, L, H, V
    o = _result(pk, np.zeros((2, 2)), np.eye(2))
    assert abs(abs(getsijmn(2, 1, R, R, o)) - 1) < 1e-15
    assert abs(getsijmn(2, 1, L, R, o)) < 1e-15
    assert abs(abs(getsijmn(2, 1, H, H, o)) - 1) < 1e-15
    assert abs(getsijmn(2, 1, V, H, o)) < 1e-15


@pytest.mark.parametrize("phi", [0.0, 30.0, 200.0])
def test_pec_reflection_flips_handedness(pk, phi):
    """At normal incidence a perfect conductor (TE and TM reflection -1) reflects
    RHCP as LHCP."""
    from pssfss_amd.outputs import getsijmn, R, L
    o = _result(pk, -np.eye(2), np.zeros((2, 2)), 0.0, phi)
    assert abs(getsijmn(1, 1, R, R, o)) < 1e-12
    assert abs(abs(getsijmn(1, 1, L, R, o)) - 1) < 1e-12


@pytest.mark.parametrize("theta", [10.0, 25.0, 40.0])
def test_oblique_basis_scaling_of_the_reference(pk, theta):
    """src/Outputs.jl:159-200 takes the TM field vector of a scattered wave as
    t2 + tan(theta) z (norm 1/cos(theta)) while the incident conversion divides the
    TM projection by cos(theta): in the plane of incidence (phi = 0) free-space
    transmission is diag(1/cos, 1) in H/V and PEC reflection of RHCP has
    co-polar part (1 - 1/cos)/2 and cross-polar part -(1 + 1/cos)/2."""
    from pssfss_amd.outputs import getsijmn, R, L, H, V
    ct = math.cos(math.radians(theta))
    o = _result(pk, np.zeros((2, 2)), np.eye(2), theta, 0.0)
    assert abs(getsijmn(2, 1, H, H, o) - 1 / ct) < 1e-12
    assert abs(getsijmn(2, 1, V, V, o) - 1) < 1e-12
    assert abs(getsijmn(2, 1, V, H, o)) < 1e-12
    p = _result(pk, -np.eye(2), np.zeros((2, 2)), theta, 0.0)
    assert abs(getsijmn(1, 1, R, R, p) - (1 - 1 / ct) / 2) < 1e-12
    assert abs(getsijmn(1, 1, L, R, p) + (1 + 1 / ct) / 2) < 1e-12


@pytest.mark.parametrize("delta", [0.0, 30.0, 90.0, 135.0])
def test_axial_ratio_matches_ellipse(pk, delta):
    """Wave plate at 45 degrees to the TE/TM axes, retardation delta: the
    transmitted field of an H-polarised wave has the ellipse's axial ratio."""
    from pssfss_amd.outputs import ardb, getsijmn, H, TE, TM
    c = s = math.sqrt(0.5)
    rot = np.array([[c, -s], [s, c]])
    s21 = rot @ np.diag([1.0, np.exp(1j * math.radians(delta))]) @ rot.T
    o = _result(pk, np.zeros((2, 2)), s21)
    e = np.array([getsijmn(2, 1, TE, H, o), getsijmn(2, 1, TM, H, o)])
    want = _ellipse_ar_db(e)
    got = ardb(2, 1, H)(o)
    if math.isinf(want):
        assert math.isinf(got)
    else:
        assert abs(got - want) < 1e-9
    if delta == 90.0:
        assert abs(got) < 1e-9          # quarter-wave plate: circular


def test_named_outputs_and_steering(pk):
    from pssfss_amd.outputs import outputs, extract_result, theta_phi
    s21 = np.diag([0.5, 0.5j])
    o = _result(pk, np.zeros((2, 2)), s21, theta=20.0, phi=35.0, fghz=12.0)
    th, ph = theta_phi(o)
    assert abs(th - 20.0) < 1e-10 and abs(ph - 35.0) < 1e-10
    ops = outputs("FGHz", "theta", "phi", "s21db(te,te)", "S21ANG(TM,TM)", "dipd21", "dil21")
    row = extract_result(o, ops)[0]
    assert row[0] == 12.0 and row[1] == 20.0 and row[2] == 35.0
    assert abs(row[3] - 20 * math.log10(0.5)) < 1e-12
    assert abs(row[4] - 90.0) < 1e-12
    assert abs(row[5] + 90.0) < 1e-12        # angle(0.5 / 0.5j)
    assert abs(row[6]) < 1e-12
    with pytest.raises(ValueError):
        outputs("s31db(te,te)")


def test_outputs_csv(pk, tmp_path):
    """The outlist CSV format of the reference (src/PSSFSS.jl:316-324, writedlm with
    ','): a header of output labels, then one row per result; complex entries as
    Julia prints them."""
    from pssfss_amd.outputs import outputs, write_outputs_csv, extract_result
    rs = [_result(pk, np.zeros((2, 2)), np.eye(2) * (0.5 + 0.25j), 0.0, 0.0, f)
          for f in (10.0, 12.5)]
    ops = outputs("FGHz", "s21(te,te)", "s21db(te,te)", "s11mag(tm,tm)")
    path = tmp_path / "out.csv"
    write_outputs_csv(str(path), rs, ops)
    lines = path.read_text().splitlines()
    assert lines[0] == ",".join(o.label for o in ops)
    assert len(lines) == 3
    f = lines[1].split(",")
    assert f[0] == "10.0" and f[1] == "0.5 + 0.25im" and f[3] == "0.0"
    assert abs(float(f[2]) - 10 * math.log10(0.5 ** 2 + 0.25 ** 2)) < 1e-12
    vals = extract_result(rs, ops)
    assert float(lines[2].split(",")[0]) == vals[1, 0].real == 12.5
