"""Residual-interaction coupling constants of the pnFAM functional.

Restates pnfam_interaction.f90:90-260 (init_interaction) and :267-715
(load_ph_interaction): Skyrme t/x parameters -> isovector EDF couplings
(E. Perlinska et al., PRC 69, 014316 (2004)), gauge-invariance fixes for C^T and
C^F, pairing couplings with the HFBTHO density dependence, namelist overrides,
and the density-dependent couplings crho(r), cs(r), cpair(r), cspair(r) on the
quadrature grid.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

HUGE = np.finfo(np.float64).max

# name -> (t0, t1, t2, t3, x0, x1, x2, x3, b4, b4p, alpha, tte, tto, uses_j2)
# (pnfam_interaction.f90:344-660; w=huge means b4/b4p are given directly)
SKYRME = {
    "SIII": (-1128.75, 395.0, -95.0, 14000.0, 0.45, 0.0, 0.0, 1.0, 60.0, 60.0, 1.0, 0.0, 0.0, False),
    "SGII": (-2645.0, 340.0, -41.9, 15595.0, 0.09, -0.0588, 1.425, 0.06044, 52.5, 52.5, 1.0 / 6.0,
             0.0, 0.0, False),
    "SKM*": (-2645.0, 410.0, -135.0, 15595.0, 0.09, 0.0, 0.0, 0.0, 65.0, 65.0, 1.0 / 6.0, 0.0, 0.0, False),
    "SKO": (-2103.653, 303.352, 791.674, 13553.252, -0.210701, -2.810752, -1.461595, -0.429881,
            176.578, -198.7490, 0.25, 0.0, 0.0, False),
    "SKOP": (-2099.419, 301.531, 154.781, 13526.464, -0.029503, -1.325732, -2.323439, -0.147404,
             143.895, -82.8888, 0.25, 0.0, 0.0, True),
    "SLY4": (-2488.913, 486.818, -546.395, 13777.0, 0.834, -0.344, -1.0, 1.354, 61.5, 61.5, 1.0 / 6.0,
             0.0, 0.0, False),
    "UNEDF0": (-1883.68781034247695, 277.500212238931113, 608.430905591534383, 13901.9483446343256,
               0.00974374651612197606, -1.77784394560870229, -1.67699034528797575,
               -0.38079041463310026, 125.161, -91.2604, 0.321955989588264435, 0.0, 0.0, False),
    "UNEDF1": (-2078.3280232556408, 239.400812041522045, 1575.11954189757989, 14263.646247077595,
               0.0537569206858470316, -5.07723238187693671, -1.36650561394298609,
               -0.162491168089956339, 38.3680720616682009, 71.3165222295833985,
               0.2700180115027076, 0.0, 0.0, False),
    "UNEDF2": (-1735.45685190850349, 262.814456235313969, 1183.31263243304011, 12293.2432941125371,
               0.172247227624101162, -3.76692685046941023, -1.38349812677880979,
               0.0528642727436126059, 25.658667730648304, 77.300389370271006,
               0.351455132555483607, -240.140598215558754, -266.930662318713644, True),
}
ALIASES = {"S3": "SIII", "SG2": "SGII", "SKO'": "SKOP", "UNE0": "UNEDF0", "UNE1": "UNEDF1",
           "UNE2": "UNEDF2"}


@dataclass
class Couplings:
    """Scalar couplings plus the four density-dependent ones on the grid."""
    skip_residual: bool
    crho: np.ndarray
    cs: np.ndarray
    cpair: np.ndarray
    cspair: np.ndarray
    cdrho: float = 0.0
    ctau: float = 0.0
    ctj0: float = 0.0
    ctj1: float = 0.0
    ctj2: float = 0.0
    crdj: float = 0.0
    cds: float = 0.0
    ct: float = 0.0
    cj: float = 0.0
    cgs: float = 0.0
    cf: float = 0.0
    csdj: float = 0.0
    # constants for printing / checks
    cr0: float = 0.0
    crr: float = 0.0
    cs0: float = 0.0
    csr: float = 0.0
    sigma: float = 0.0
    cpair0: float = 0.0
    cpairr: float = 0.0
    cspair0: float = 0.0
    cspairr: float = 0.0
    name: str = ""

    def scalars(self):
        """The 12 scalar couplings in the order the device kernels take them."""
        return np.array([self.cdrho, self.ctau, self.ctj0, self.ctj1, self.ctj2, self.crdj,
                         self.cds, self.ct, self.cj, self.cgs, self.cf, self.csdj], dtype=np.float64)


def _unset(v):
    return v is None or (isinstance(v, float) and abs(v - HUGE) <= np.finfo(float).eps) or v == ""


def init_interaction(name, rho_total, hfb_cpair, hfb_alpha_pair, rho_c, *,
                     require_gauge_invariance=True, force_j2_terms=False,
                     vpair_t0=None, vpair_t1=None, vpair0=None, vpair1=None,
                     override_cs0=None, override_csr=None, override_cds=None, override_ct=None,
                     override_cf=None, override_cgs=None, override_cj=None, override_csdj=None):
    nghl = rho_total.shape[0]
    key = name.strip().upper()
    key = ALIASES.get(key, key)
    if key == "NONE":
        z = np.zeros(nghl)
        return Couplings(skip_residual=True, crho=z, cs=z.copy(), cpair=z.copy(), cspair=z.copy(),
                         name="NONE")
    if key not in SKYRME:
        raise ValueError(f"interaction {name!r} not found")
    t0, t1, t2, t3, x0, x1, x2, x3, b4, b4p, alpha, tte, tto, uses_j2 = SKYRME[key]
    c = Couplings(skip_residual=False, crho=None, cs=None, cpair=None, cspair=None, name=key)
    c.cr0 = -1.0 / 8.0 * t0 * (2.0 * x0 + 1.0)
    c.crr = -1.0 / 48.0 * t3 * (2.0 * x3 + 1.0)
    c.cs0 = -1.0 / 8.0 * t0
    c.csr = -1.0 / 48.0 * t3
    c.sigma = alpha
    c.cdrho = 3.0 / 32.0 * t1 * (x1 + 0.5) + 1.0 / 32.0 * t2 * (x2 + 0.5)
    c.ctau = -1.0 / 8.0 * t1 * (x1 + 0.5) + 1.0 / 8.0 * t2 * (x2 + 0.5)
    c.ctj0 = 1.0 / 48.0 * (t1 - t2 + 10.0 * tte - 10.0 * tto)
    c.ctj1 = 1.0 / 32.0 * (t1 - t2 - 5.0 * tte + 5.0 * tto)
    c.ctj2 = 1.0 / 16.0 * (t1 - t2 + tte - tto)
    c.crdj = -0.5 * b4p
    c.cds = 1.0 / 64.0 * (3.0 * t1 + t2 - 6.0 * tte - 2.0 * tto)
    c.ct = -1.0 / 16.0 * (t1 - t2 - 2.0 * tte + 2.0 * tto)
    c.cj = 1.0 / 8.0 * t1 * (x1 + 0.5) - 1.0 / 8.0 * t2 * (x2 + 0.5)
    c.cgs = -3.0 / 32.0 * (3.0 * tte + tto)
    c.cf = -3.0 / 8.0 * (tte - tto)
    c.csdj = -0.5 * b4p
    # single tensor coupling as in HFBTHO
    c.ctj0 = 2.0 / 3.0 * c.ctj1
    c.ctj2 = 2.0 * c.ctj1
    if not uses_j2 and not force_j2_terms:
        c.ctj0 = c.ctj1 = c.ctj2 = 0.0
    if require_gauge_invariance:
        c.cf = 2.0 * c.ctj1 - c.ctj2
        c.ct = -2.0 * c.ctj1 + 0.5 * c.cf
    # pairing strengths
    if _unset(vpair_t0):
        vpair_t0 = 0.0 if _unset(vpair0) else 2.0 * vpair0
    elif not _unset(vpair0):
        raise ValueError("cannot set both vpair0 and vpair_t0")
    if _unset(vpair_t1):
        vpair_t1 = 0.5 * float(np.sum(hfb_cpair)) if _unset(vpair1) else vpair1
    elif not _unset(vpair1):
        raise ValueError("cannot set both vpair1 and vpair_t1")
    sigma_pair = 1.0
    c.cpair0 = vpair_t1 / 8.0
    c.cspair0 = vpair_t0 / 8.0
    c.cpairr = c.cpair0 * (-1.0 * hfb_alpha_pair / rho_c ** sigma_pair)
    c.cspairr = c.cspair0 * (-1.0 * hfb_alpha_pair / rho_c ** sigma_pair)
    for attr, val in (("cs0", override_cs0), ("csr", override_csr), ("cds", override_cds),
                      ("ct", override_ct), ("cf", override_cf), ("cgs", override_cgs),
                      ("cj", override_cj), ("csdj", override_csdj)):
        if not _unset(val):
            setattr(c, attr, float(val))
    rho = rho_total
    c.crho = c.cr0 + c.crr * rho ** c.sigma
    c.cs = c.cs0 + c.csr * rho ** c.sigma
    c.cpair = c.cpair0 + c.cpairr * rho ** sigma_pair
    c.cspair = c.cspair0 + c.cspairr * rho ** sigma_pair
    return c
