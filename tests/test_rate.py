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


# ---- first-forbidden shape factors and rates (all-operator S40 run) ----------

def load_all_operator_run():
    """Archived contour tables of the all-operator S40 run as full-contour
    strengths {(op, K): S} and cross-terms {(label, K): G}, with EQRPA_max (from
    the contour), the HFB Q value (EQRPA_max - E_gs) and the run's summary."""
    with open(os.path.join(GOLDEN, "S40_All", "rates.json")) as fh:
        d = json.load(fh)
    strengths, cross = {}, {}
    for name, tab in d["contour"].items():
        op, k = name.split("-K")
        cols, a = tab["columns"], np.array(tab["rows"], dtype=float)
        strengths[(op, int(k))] = a[:, cols.index("Re(Strength)")] + 1j * a[:, cols.index("Im(Strength)")]
        for c in cols:
            if c.startswith("Re(") and "x" in c:
                lbl = c[3:-1]
                cross[(lbl, int(k))] = a[:, cols.index(c)] + 1j * a[:, cols.index("Im(%s)" % lbl)]
    a = np.array(d["contour"]["GT-K0"]["rows"], dtype=float)
    z = a[:, 1] + 1j * a[:, 2]
    emax = 2.0 * float(np.mean((z / (1.0 + np.exp(1j * a[:, 0]))).real))
    return d, strengths, cross, emax, emax - d["E_gs"]


def _archived_phase(d):
    pr, pi = np.array(d["phasespace_re"]["rows"], dtype=float), np.array(d["phasespace_im"]["rows"], dtype=float)
    return {n: pr[:, 3 + n] + 1j * pi[:, 3 + n] for n in range(1, 7)}


def test_ff_phase_space_matches_archived():
    """f1..f6 of the all-operator run; f6 carries lambda_2 = F1/F0 times the
    Coulomb factor, i.e. the generalised Fermi function F1."""
    d, _, _, emax, _ = load_all_operator_run()
    s = d["summary"]
    ref = _archived_phase(d)
    pr = np.array(d["phasespace_re"]["rows"], dtype=float)
    w0 = (emax - (pr[:, 0] + 1j * pr[:, 1])) / R.MEC2 + 1.0
    for n in range(1, 7):
        f = R.phase_space_function(n, s["Zf"], s["A"], emax, 0.0)(w0)
        assert np.max(np.abs(f - ref[n])) < 1e-9 * np.max(np.abs(ref[n])), n


def test_generalised_fermi_function_limits():
    """Z = 0: F_{k-1} = (k (2k-1)!!)^2 4^k Gamma(k)^2 / Gamma(2k+1)^2 = 1 for
    k = 1, 2 (36 * 16 / 24^2 for k = 2), so lambda_2 = (2 + 2) / (2 * 2) = 1; at
    and below threshold lambda_2 is zero."""
    w = np.array([1.2, 3.0, 10.0])
    assert np.allclose(R.fermi_fk(2, 0.0, 100, w), 1.0, rtol=0, atol=1e-13)
    assert np.allclose(R.lambda_2(0.0, 100, w), 1.0, rtol=0, atol=1e-13)
    assert np.all(R.lambda_2(30.0, 100, np.array([0.5, 1.0])) == 0.0)


def test_ff_shape_factors_match_archived():
    """Every beta.out channel's shape factor on the contour, from the archived
    strengths, cross-terms and phase space: the prefactors, the Coulomb x+/-
    and each J term pinned by its own channel ((J, K) for J = 0, 1, 2)."""
    from pynfam_amd import ff_rate as FF
    d, S, G, emax, q = load_all_operator_run()
    s = d["summary"]
    _, sf, _ = FF.beta_rates(S, G, s["Zf"], s["A"], emax, q, lam=abs(s["gA"] / s["gV"]), phase=_archived_phase(d))
    cols = d["shapefactor_re"]["columns"]
    ar, ai = np.array(d["shapefactor_re"]["rows"], dtype=float), np.array(d["shapefactor_im"]["rows"], dtype=float)
    assert set(FF.CHANNELS) == set(cols) - {"Re(EQRPA)", "Im(EQRPA)"}
    for name in FF.CHANNELS:
        ref = ar[:, cols.index(name)] + 1j * ai[:, cols.index(name)]
        assert np.max(np.abs(sf[name] - ref)) < 1e-13 * np.max(np.abs(ref)), name


def test_ff_rates_match_archived():
    """All 19 rates of beta.out: 1e-10 relative with the archived phase space;
    with this module's RATINT phase space (1e-10 of max |f| off near the
    endpoint) the closed-contour cancellation leaves 1.2e-7 on the small Fermi
    rate and <= 2e-8 elsewhere: 1e-6 required."""
    from pynfam_amd import ff_rate as FF
    d, S, G, emax, q = load_all_operator_run()
    s = d["summary"]
    lam = abs(s["gA"] / s["gV"])
    ref = d["rates"]
    assert set(ref) == set(FF.CHANNELS)
    r1, _, _ = FF.beta_rates(S, G, s["Zf"], s["A"], emax, q, lam=lam, phase=_archived_phase(d))
    r2, _, _ = FF.beta_rates(S, G, s["Zf"], s["A"], emax, q, lam=lam)
    for name, v in ref.items():
        assert abs(r1[name] - v) < 1e-10 * abs(v), (name, r1[name], v)
        assert abs(r2[name] - v) < 1e-6 * abs(v), (name, r2[name], v)
    # the channel sums are consistent: per K, per J and per (J, K) all add up to the total
    tf = r1["Total-Forbidden"]
    for group in (["Forbidden-K=%d" % k for k in range(3)], ["Forbidden-J=%d" % j for j in range(3)],
                  [c for c in FF.CHANNELS if c.startswith("Forbidden-(J,K)")]):
        assert abs(sum(r1[c] for c in group) - tf) < 1e-12 * tf, group
    assert abs(r1["Total-Allowed"] + tf - r1["Total"]) < 1e-12 * r1["Total"]


def test_ff_cross_terms_follow_the_conjugate_fill():
    """The archived second half of every strength and cross-term column is the
    conjugate fill of the first half, so the engine computes only the first."""
    _, S, G, _, _ = load_all_operator_run()
    for v in list(S.values()) + list(G.values()):
        n = len(v)
        assert np.array_equal(R.fill_conjugate(v[:(n + 1) // 2], n), v)


def test_ff_rejects_unknown_operators():
    """Operators without a shape-factor term (the I-function set, two-body
    currents) are refused rather than dropped."""
    from pynfam_amd import ff_rate as FF
    with pytest.raises(ValueError):
        FF.strength_products({("RS0I", 0): np.ones(4)}, {}, 4)
