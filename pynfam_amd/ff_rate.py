"""Allowed and first-forbidden beta-minus shape factors and rates.

pynfam turns the strengths of the one-body operators on a closed contour into
beta-decay rates (pynfam/strength/shape_factor.py).  This module restates that
step for beta-minus, so that a rate can be formed from strengths and cross-terms
computed on either precision path of the engine:

- the dimensionful operator prefactors and the products that enter the shape
  factor: b_op,K = pre_op,K^2 S_op,K, and for the four interfering pairs
  b_a_b,K = pre_a,K pre_b,K G_axb,K (shape_factor.py:527-646);
- the shape-factor terms of rank J = 0, 1, 2 with the phase-space functions
  f1..f6 and the Coulomb quantities x+/- = W0_max/3 +/- alpha Z / 2R
  (Behrens-Buhring form, shape_factor.py:649-783);
- the 19 rate channels of beta.out: totals, allowed Fermi and GT per K, and
  the forbidden sums per K, per J and per (J, K) (:786-839), each integrated
  on the closed contour as the allowed rates are (:956-1029, 1160-1196).

The shape factor is linear in the strengths, so every term is formed on the
whole contour and summed channel by channel.  Beta-plus and electron capture
(other Coulomb signs and endpoints) are not covered.
"""
from __future__ import annotations

import numpy as np

from . import rate as R

MN = 939.0          # nucleon mass [MeV], pynfam/config.py:228
SQ2, SQ3 = np.sqrt(2.0), np.sqrt(3.0)

OPS = ("F", "GT", "P", "R", "PS0", "RS0", "RS1", "RS2")
# interfering pairs: (key, operator whose contour table holds it, cross-term label, partner)
CROSS = (("RS0_PS0", "RS0", "RS0xPS0", "PS0"), ("R_RS1", "R", "RxRS1", "RS1"),
         ("P_RS1", "P", "PxRS1", "RS1"), ("R_P", "R", "RxP", "P"))

TERMS_J = {0: ("J0_R1", "J0_R2"),
           1: ("J1_R1", "J1_R2", "J1_R3", "J1_R4", "J1_R5", "J1_R6"),
           2: ("J2_R5", "J2_R6")}
FF_TERMS = TERMS_J[0] + TERMS_J[1] + TERMS_J[2]


def prefactor(op, K):
    """Dimensionful prefactor of operator op with projection K: T_K = 1 for
    K = 0 and sqrt(2) for K > 0; R and RS in units of the electron Compton
    wavelength, P and PS0 with hbar c m_e c^2 / M_N (shape_factor.py:548-570)."""
    tk = 1.0 if K == 0 else SQ2
    pmom = R.HBAR_MEC * R.MEC2 / MN
    table = {"F": tk, "GT": tk, "P": -tk * pmom, "R": tk * SQ3 / R.HBAR_MEC, "PS0": -pmom,
             "RS0": 1.0 / R.HBAR_MEC, "RS1": -tk / R.HBAR_MEC, "RS2": tk / R.HBAR_MEC}
    return table[op]


def strength_products(strengths, crossterms, npts):
    """b[key][K] on the contour from strengths {(op, K): S} and cross-terms
    {(label, K): G}, both on the whole contour.  Operator-K pairs that were not
    computed contribute zero, as in the reference."""
    b = {key: {K: np.zeros(npts, dtype=complex) for K in range(3)} for key in OPS + tuple(c[0] for c in CROSS)}
    for (op, K), s in strengths.items():
        if op not in OPS or not 0 <= K <= 2:
            raise ValueError(f"no shape-factor term for operator {op} with K={K}")
        b[op][K] = prefactor(op, K) ** 2 * np.asarray(s, dtype=complex)
    for key, op, label, partner in CROSS:
        for K in range(3):
            if (label, K) in crossterms:
                b[key][K] = prefactor(op, K) * prefactor(partner, K) * np.asarray(crossterms[(label, K)], dtype=complex)
    return b


def coulomb_x(zd, a, w0max):
    """x+ and x- = W0_max / 3 +/- alpha Z / 2R (I(1,1,1,1; r) ~ 3/2), with R the
    nuclear radius in units of the electron Compton wavelength."""
    rad = R.R0 * a ** (1.0 / 3.0) / R.HBAR_MEC
    c = 0.5 * R.ALPHA * zd / rad
    return w0max / 3.0 + c, w0max / 3.0 - c


def shape_factor_terms(b, f, zd, a, w0max, lam):
    """The shape-factor terms {name: {K: C}} on the contour: allowed F and GT,
    and the first-forbidden terms J0_R1..J2_R6 (beta-minus signs).  f is
    {n: f_n on the contour}, lam = |gA / gV|."""
    xp, xm = coulomb_x(zd, a, w0max)
    g1 = R.gamma_k(1, zd)
    l2 = lam * lam
    C = {}
    for K in range(3):
        S = {key: v[K] for key, v in b.items()}
        rs0, ps0, r, p, rs1, rs2 = S["RS0"], S["PS0"], S["R"], S["P"], S["RS1"], S["RS2"]
        x_rs0ps0, x_rrs1, x_prs1, x_rp = S["RS0_PS0"], S["R_RS1"], S["P_RS1"], S["R_P"]
        t = {
            "F": S["F"] * f[2],
            "GT": l2 * S["GT"] * f[2],
            # rank 0: RS0 and PS0
            "J0_R1": -2.0 / 3.0 * l2 * (xp * rs0 + x_rs0ps0) * f[1],
            "J0_R2": l2 * ((xp * xp + 1.0 / 9.0) * rs0 + ps0 + 2.0 * xp * x_rs0ps0) * f[2],
            # rank 1: R, P and RS1 with their three interference terms
            "J1_R1": -2.0 / 9.0 * (xp * r - 2.0 * l2 * xm * rs1 + lam * SQ2 * (xp - xm) * x_rrs1
                                   - SQ3 * x_rp - lam * np.sqrt(6.0) * x_prs1) * f[1],
            "J1_R2": (p + xp * xp / 3.0 * r + 2.0 / 3.0 * l2 * xm * xm * rs1
                      + (r + 2.0 * l2 * rs1 + 2.0 * SQ2 * lam * x_rrs1) / 27.0
                      + np.sqrt(2.0 / 3.0) * (2.0 * lam * xm * x_prs1 - SQ2 * xp * x_rp
                                              - 2.0 / SQ3 * lam * xm * xp * x_rrs1)
                      - 8.0 / 27.0 * (l2 * rs1 + lam / SQ2 * x_rrs1) * g1) * f[2],
            "J1_R3": 4.0 / 3.0 * (SQ2 / 3.0 * lam * xp * x_rrs1 - 2.0 / 3.0 * l2 * xm * rs1
                                  - np.sqrt(2.0 / 3.0) * lam * x_prs1) * f[3],
            "J1_R4": 8.0 / 27.0 * l2 * rs1 * f[4],
            "J1_R5": (2.0 * r + l2 * rs1 + 2.0 * SQ2 * lam * x_rrs1) / 27.0 * f[5],
            "J1_R6": (2.0 * r + l2 * rs1 - 2.0 * SQ2 * lam * x_rrs1) / 27.0 * f[6],
            # rank 2: RS2
            "J2_R5": l2 / 9.0 * rs2 * f[5],
            "J2_R6": l2 / 9.0 * rs2 * f[6],
        }
        for name, v in t.items():
            C.setdefault(name, {})[K] = v
    return C


def _channels():
    """beta.out channel -> (terms, K values) summed into it (shape_factor.py:786-839)."""
    allowed = ("F", "GT")
    ch = {
        "Total": (allowed + FF_TERMS, (0, 1, 2)),
        "Total-Allowed": (allowed, (0, 1)),
        "Total-GT": (("GT",), (0, 1)),
        "Total-Forbidden": (FF_TERMS, (0, 1, 2)),
        "Allowed-Fermi": (("F",), (0,)),
        "Allowed-GT_K=0": (("GT",), (0,)),
        "Allowed-GT_K=1": (("GT",), (1,)),
    }
    for K in range(3):
        ch["Forbidden-K=%d" % K] = (FF_TERMS, (K,))
    for J in range(3):
        ch["Forbidden-J=%d" % J] = (TERMS_J[J], (0, 1, 2))
    for J, K in ((0, 0), (1, 0), (1, 1), (2, 0), (2, 1), (2, 2)):
        ch["Forbidden-(J,K)=(%d,%d)" % (J, K)] = (TERMS_J[J], (K,))
    return ch


CHANNELS = _channels()


def shape_factor_channels(C, npts):
    """Shape factor of every beta.out channel on the contour."""
    out = {}
    for name, (terms, ks) in CHANNELS.items():
        acc = np.zeros(npts, dtype=complex)
        for t in terms:
            for K in ks:
                acc = acc + C[t][K]
        out[name] = acc
    return out


def contour_rate(contour, sf):
    """Rate [1/s] of a shape factor on the closed contour:
    ln2 / kappa Im(-i pi (1 / 2 pi i) sum_k C_k dz/dt_k w_k)."""
    cint = np.sum(np.asarray(sf) * contour["dzdt"] * contour["weights"]) / (2j * np.pi)
    return float(np.log(2.0) / R.KAPPA * np.imag(-1j * np.pi * cint))


def beta_rates(strengths, crossterms, zd, a, eqrpa_max, q_value, emin=0.0, lam=abs(R.GA / R.GV), phase=None):
    """All 19 beta.out channels from full-contour strengths {(op, K): S} and
    cross-terms {(label, K): G} on the CIRCLE contour through emin and
    eqrpa_max.  q_value is the HFB Q value (EQRPA_max - E_gs); phase, if
    given, is {n: f_n on the contour} instead of the RATINT continuation.
    Returns (rates, shape factors per channel, contour)."""
    npts = len(next(iter(strengths.values())))
    ctr = R.circle_contour(emin, eqrpa_max, npts)
    if phase is None:
        w0 = (eqrpa_max - ctr["z"]) / R.MEC2 + 1.0
        phase = {n: R.phase_space_function(n, zd, a, eqrpa_max, emin)(w0) for n in range(1, 7)}
    w0max = 1.0 + q_value / R.MEC2
    b = strength_products(strengths, crossterms, npts)
    C = shape_factor_terms(b, phase, zd, a, w0max, lam)
    sf = shape_factor_channels(C, npts)
    rates = {name: contour_rate(ctr, v) for name, v in sf.items()}
    return rates, sf, ctr
