"""Contour, allowed beta-minus phase space and rates from pnFAM strengths.

The consumer side of the pnFAM iteration: pynfam integrates S(w) on a closed
contour against the lepton phase space to get the beta-decay rate
(pynfam/strength/contour.py, phase_space.py, shape_factor.py).  This module
restates the parts the allowed Gamow-Teller and Fermi rates need, so that a rate
can be formed from strengths computed on either precision path:

- the CIRCLE contour with Gauss-Legendre nodes in the angle and the half-contour
  symmetry fill S(w*) = S(w)* (contour.py:212-281, fam_strength.py:295-330);
- the phase-space integrals f_n(W0) on the real axis (Gauss-Legendre in the
  electron energy, phase_space.py:253-356), with the point-charge Fermi function
  F0 L0 (:845-982), continued into the complex plane by Thiele's rational
  interpolation on Chebyshev-Gauss nodes (RATINT, :103-148, 668-708);
- the closed-contour quadrature and rate (shape_factor.py:956-1029, 1160-1196):
  rate = ln2 / kappa * Im(-i pi (1 / 2 pi i) sum_k C_k dz/dt_k w_k), with the GT
  shape factor C = lambda^2 T_K^2 S f_2 (T_0 = 1, T_1 = sqrt 2; :548-560, 676-677).

Only beta-minus without screening is covered (the reference's default for it).
"""
from __future__ import annotations

import numpy as np
from scipy.special import loggamma

# physical constants as pynfam/config.py:221-228 sets them
ALPHA = 7.2973525698e-03
HBAR_MEC = 386.15926800      # hbar c / m_e c^2 [fm]
MEC2 = 0.510998928           # m_e c^2 [MeV]
KAPPA = 6147.0               # [s]
DMASS_NH = 0.78227           # n - H mass difference [MeV]
R0 = 1.2                     # [fm]
GA = -1.27
GV = 1.0
PSI_GLPTS = 15               # pynfam/config.py:201-211 defaults
RATINT_PTS = 20
_EPS = np.finfo(float).eps


# ---- contour ---------------------------------------------------------------
def circle_contour(emin, emax, npts, theta0=np.pi):
    """Gauss-Legendre nodes in the angle on [theta0, theta0 + 2 pi] of the circle
    through emin and emax.  Returns z, dz/dtheta, theta, weights and the number
    of leading points to compute (the rest follow by conjugation when the start
    angle is a multiple of pi)."""
    x, wts = np.polynomial.legendre.leggauss(npts)
    half = np.pi
    theta = x * half + theta0 + half
    wts = wts * half
    c = 0.5 * (emin + emax)
    r = emax - c
    z = c + r * np.exp(1j * theta)
    dzdt = 1j * (z - c)
    ncomp = (npts + 1) // 2 if abs(np.mod(theta0, np.pi)) <= 1e-10 else npts
    return {"z": z, "dzdt": dzdt, "theta": theta, "weights": wts, "n_compute": ncomp}


def fill_conjugate(s_half, npts):
    """Strengths on the whole contour from the first (npts+1)//2 points:
    S(w*) = S(w)*, the mirrored half in reverse order (the middle point of an
    odd contour is not repeated)."""
    s_half = np.asarray(s_half, dtype=complex)
    mirror = np.conj(s_half[::-1])
    if 2 * len(s_half) != npts:
        mirror = mirror[1:]
    return np.concatenate([s_half, mirror])


# ---- phase space ------------------------------------------------------------
def gamma_k(k, zd):
    return np.sqrt(float(k) ** 2 - (ALPHA * zd) ** 2)


def coulomb_l0(zd):
    """Simplest L0 without screening."""
    return 0.5 * (1.0 + gamma_k(1, zd))


def _odd_factorial2(n):
    """(2k-1)!! for odd n = 2k-1 >= 1."""
    out = 1
    for j in range(n, 0, -2):
        out *= j
    return out


def fermi_fk(k, zd, a, w):
    """Generalised point-charge Fermi function F_{k-1} of electron energy w
    (units of m_e c^2), without L0 (phase_space.py:881-982):
        (k (2k-1)!!)^2 4^k (2 p R)^(2 (gamma_k - k)) e^(pi y) |Gamma(gamma_k + i y)|^2
        / Gamma(2 gamma_k + 1)^2,  y = alpha Z w / p.
    Evaluated through its logarithm: exp(pi y) and |Gamma(g + iy)|^2 over- and
    underflow separately near threshold.  Zero at and below w = 1."""
    w = np.asarray(w, dtype=float)
    out = np.zeros_like(w)
    ok = ~((np.abs(w - 1.0) < _EPS) | (w < 1.0))
    if not np.any(ok):
        return out
    wk = w[ok]
    p = np.sqrt(wk * wk - 1.0)
    radius = R0 / HBAR_MEC * a ** (1.0 / 3.0)
    g = gamma_k(k, zd)
    y = ALPHA * zd * wk / p
    lg = loggamma(g + 1j * y)
    ln_f = (2.0 * np.log(float(k * _odd_factorial2(2 * k - 1))) + k * np.log(4.0)
            + 2.0 * (g - k) * np.log(2.0 * p * radius) + np.pi * y + 2.0 * lg.real
            - 2.0 * loggamma(2.0 * g + 1.0).real)
    out[ok] = np.exp(ln_f)
    return out


def fermi_f0(zd, a, w):
    """Point-charge Fermi function F0 (k = 1 of fermi_fk)."""
    return fermi_fk(1, zd, a, w)


def lambda_2(zd, a, w):
    """Coulomb function lambda_2 = (2 + gamma_2) / (2 (1 + gamma_1)) F1 / F0
    (phase_space.py:785-826); zero where F0 vanishes."""
    w = np.asarray(w, dtype=float)
    f0, f1 = fermi_fk(1, zd, a, w), fermi_fk(2, zd, a, w)
    out = np.zeros_like(w)
    ok = f0 > 0.0
    out[ok] = (2.0 + gamma_k(2, zd)) / (2.0 * (1.0 + gamma_k(1, zd))) * f1[ok] / f0[ok]
    return out


def psi_integrand(wx, n, zd, a, w0):
    """Integrand of f_n (n = 1..6) for beta-minus at electron energy wx and
    endpoint w0 (phase_space.py:253-313)."""
    wx = np.maximum(np.asarray(wx, dtype=float), 1.0)
    px = np.sqrt(wx * wx - 1.0)
    pv = w0 - wx
    if n == 6:
        gn = wx * px * px * lambda_2(zd, a, wx)
    else:
        gn = {1: gamma_k(1, zd), 2: wx, 3: wx * wx, 4: wx ** 3, 5: wx * pv * pv}[n]
    f = px * pv * pv * gn * fermi_f0(zd, a, wx) * coulomb_l0(zd)
    f[~np.isfinite(f)] = 0.0
    return f


def phase_space_real(n, zd, a, w0, glpts=PSI_GLPTS):
    """f_n(W0) for real W0 >= 1: Gauss-Legendre on [1, W0]."""
    x, wts = np.polynomial.legendre.leggauss(glpts)

    def one(w0s):
        h = 0.5 * (w0s - 1.0)
        wx = x * h + 1.0 + h
        return np.sum(psi_integrand(wx, n, zd, a, w0s) * wts * h)

    return np.vectorize(one)(np.asarray(w0, dtype=float))


def thiele(x, y):
    """Thiele continued-fraction interpolant through (x_i, y_i); returns a
    function of real or complex argument."""
    x = np.asarray(x, dtype=float)
    n = len(x)
    # reciprocal differences rho[j][i] built column by column
    rho = [list(np.asarray(y, dtype=float))]
    col1 = [(x[j] - x[j + 1]) / (rho[0][j] - rho[0][j + 1] + 1e-15) for j in range(n - 1)]
    rho.append(col1)
    for i in range(2, n):
        prev, prev2 = rho[i - 1], rho[i - 2]
        rho.append([(x[j] - x[j + i]) / (prev[j] - prev[j + 1]) + prev2[j + 1] for j in range(n - i)])
    top = [rho[i][0] for i in range(n)]

    def f(z):
        acc = 0.0
        for i in range(n - 1, 1, -1):
            acc = (z - x[i - 1]) / (top[i] - top[i - 2] + acc)
        return top[0] + (z - x[0]) / (top[1] + acc)

    return f


def phase_space_function(n, zd, a, eqrpa_max, eqrpa_min, npts=RATINT_PTS):
    """f_n as a function of complex W0: Thiele interpolation of the real-axis
    integrals on Chebyshev-Gauss nodes of [1, 1 + (EQRPA_max - EQRPA_min)/m_e c^2]."""
    lo, hi = 1.0, 1.0 + (eqrpa_max - eqrpa_min) / MEC2
    xc, _ = np.polynomial.chebyshev.chebgauss(npts)
    w0 = 0.5 * (hi - lo) * xc + 0.5 * (hi + lo)
    return thiele(w0, phase_space_real(n, zd, a, w0))


# ---- rates ------------------------------------------------------------------
def eqrpa_max_beta_minus(lambda_n, lambda_p):
    """Largest QRPA energy open to beta-minus decay: Delta M_nH + lambda_n - lambda_p
    (pynfam/fortran/hfbtho_run.py:479-530)."""
    return DMASS_NH + lambda_n - lambda_p


def gt_rates(contour, strengths, f2):
    """Allowed GT rates per K from full-contour strengths {K: S} and the phase
    space f2 on the contour; returns {K: rate} and the total (K and -K counted
    through T_K^2)."""
    lam2 = (GA / GV) ** 2
    rates, cond = {}, {}
    for K, s in strengths.items():
        t2 = 1.0 if K == 0 else 2.0
        c = lam2 * t2 * np.asarray(s) * f2
        terms = c * contour["dzdt"] * contour["weights"] / (2j * np.pi)
        cint = np.sum(terms)
        rates[K] = float(np.log(2.0) / KAPPA * np.imag(-1j * np.pi * cint))
        # sum of |terms| over |rate integral|: how much the quadrature cancels
        cond[K] = float(np.sum(np.abs(np.real(terms))) / max(abs(np.real(cint)), 1e-300))
    rates["total"] = float(sum(rates[K] for K in strengths))
    rates["cancellation"] = cond
    return rates


def gt_rates_from_half(zd, a, eqrpa_max, emin, npts, s_half_by_k):
    """Contour of npts points on [emin, eqrpa_max], phase space f2 on it and the GT
    rates from strengths on the computed half."""
    ctr = circle_contour(emin, eqrpa_max, npts)
    f2fun = phase_space_function(2, zd, a, eqrpa_max, emin)
    w0 = (eqrpa_max - ctr["z"]) / MEC2 + 1.0
    f2 = f2fun(w0)
    full = {K: fill_conjugate(s, npts) for K, s in s_half_by_k.items()}
    return gt_rates(ctr, full, f2), ctr, f2
