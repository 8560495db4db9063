"""Contour, phase space and rates (pynfam_amd/rate.py) against the archived
beta-decay post-processing of the reference runs (tests/golden/*/rates.json:
the contour tables GT-K*.out, phasespace_re/im.out and beta.out)."""
import json
import os

import numpy as np
import pytest

from pynfam_amd import rate as R

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(nucleus):
    with open(os.path.join(GOLDEN, nucleus, "rates.json")) as fh:
        d = json.load(fh)

    def arr(tab):
        return np.array([[float(x) if not isinstance(x, str) else np.nan for x in r] for r in tab["rows"]])
    return d, arr


def _eqrpa_max(d, arr):
    """The EQRPA_max the archived run used, recovered from its contour
    (a circle through 0 and EQRPA_max): z = c (1 + e^{i theta}), EQRPA_max = 2c."""
    a = arr(d["GT-K0"])
    z = a[:, 1] + 1j * a[:, 2]
    c = z / (1.0 + np.exp(1j * a[:, 0]))
    return 2.0 * float(np.mean(c.real))


@pytest.mark.parametrize("nucleus", ["S40", "Gd162"])
def test_contour_matches_archived(nucleus):
    d, arr = _load(nucleus)
    emax = _eqrpa_max(d, arr)
    a = arr(d["GT-K0"])
    ctr = R.circle_contour(0.0, emax, len(a))
    assert ctr["n_compute"] == len(a) // 2
    assert np.max(np.abs(ctr["theta"] - a[:, 0])) < 1e-13
    assert np.max(np.abs(ctr["weights"] - a[:, 5])) < 1e-15
    assert np.max(np.abs(ctr["z"] - (a[:, 1] + 1j * a[:, 2]))) < 1e-12
    assert np.max(np.abs(ctr["dzdt"] - (a[:, 6] + 1j * a[:, 7]))) < 1e-12
    # the second half of the archived table is the conjugate fill of the first
    s = a[:, 3] + 1j * a[:, 4]
    assert np.array_equal(R.fill_conjugate(s[:len(s) // 2], len(s)), s)


@pytest.mark.parametrize("nucleus", ["S40", "Gd162"])
def test_phase_space_matches_archived(nucleus):
    """f1..f5 on the complex contour: RATINT continuation of the real-axis
    integrals.  Points next to the endpoint W0 = 1 carry the interpolation's
    round-off, so the bound is on the error relative to the largest value."""
    d, arr = _load(nucleus)
    s = d["summary"]
    emax = _eqrpa_max(d, arr)
    pr, pi = arr(d["phasespace_re"]), arr(d["phasespace_im"])
    z = pr[:, 0] + 1j * pr[:, 1]
    w0 = (emax - z) / R.MEC2 + 1.0
    assert np.max(np.abs(w0 - (pr[:, 2] + 1j * pr[:, 3]))) < 1e-12
    for n in (1, 2, 3, 4, 5):
        f = R.phase_space_function(n, s["Zf"], s["A"], emax, 0.0)(w0)
        ref = pr[:, 3 + n] + 1j * pi[:, 3 + n]
        assert np.max(np.abs(f - ref)) < 1e-9 * np.max(np.abs(ref)), n


def test_fermi_function_limits():
    """Z = 0: F0 = 4 |Gamma(1)|^2 / Gamma(3)^2 = 1 and L0 = 1 for every energy;
    below threshold the function is zero."""
    w = np.array([1.5, 3.0, 10.0])
    assert np.allclose(R.fermi_f0(0.0, 100, w), 1.0, rtol=0, atol=1e-14)
    assert R.coulomb_l0(0.0) == 1.0
    assert np.all(R.fermi_f0(30.0, 100, np.array([0.5, 1.0])) == 0.0)
    # f_2 at Z = 0: with many nodes the quadrature converges to the closed form of
    # int_1^W0 p w (W0 - w)^2 dw (the 15-node rule of the reference differs from
    # it by ~1e-4 because of the square-root endpoint; the archived values pin that)
    w0 = 4.0
    p0 = np.sqrt(w0 * w0 - 1.0)
    # int p w dw = p^3/3, int p w^2 dw and int p w^3 dw by parts / substitution
    i1 = p0 ** 3 / 3.0
    i2 = (w0 * p0 ** 3) / 4.0 + (w0 * p0 - np.log(w0 + p0)) / 8.0
    i3 = p0 ** 5 / 5.0 + p0 ** 3 / 3.0
    exact = w0 * w0 * i1 - 2.0 * w0 * i2 + i3
    assert abs(R.phase_space_real(2, 0.0, 100, w0, glpts=400) - exact) < 1e-8 * exact


@pytest.mark.parametrize("nucleus", ["S40", "Gd162"])
def test_rates_match_archived(nucleus):
    """GT rates per K from the archived strengths on the computed half of the
    contour and this module's phase space: beta.out to 1e-8 relative."""
    d, arr = _load(nucleus)
    s = d["summary"]
    emax = _eqrpa_max(d, arr)
    half = {}
    for K in (0, 1):
        a = arr(d["GT-K%d" % K])
        half[K] = (a[:, 3] + 1j * a[:, 4])[:len(a) // 2]
    npts = len(arr(d["GT-K0"]))
    rates, _, _ = R.gt_rates_from_half(s["Zf"], s["A"], emax, 0.0, npts, half)
    ref = d["rates"]
    for K in (0, 1):
        assert abs(rates[K] - ref["Allowed-GT_K=%d" % K]) < 1e-8 * ref["Allowed-GT_K=%d" % K], (K, rates[K])
    assert abs(rates["total"] - ref["Total-GT"]) < 1e-8 * ref["Total-GT"]
