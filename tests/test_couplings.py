"""Coupling constants against the values every archived pnFAM log prints
(pnfam_txtoutput.f90 interaction header, 9 decimals)."""
import numpy as np
import pytest

from conftest import golden_points, golden_problem

MAP = {"Crho[0]": "cr0", "Crho[r]": "crr", "Cs[0]": "cs0", "Cs[r]": "csr", "Cdr": "cdrho", "Cds": "cds",
       "Ctau": "ctau", "Cj": "cj", "CrdJ": "crdj", "Csdj": "csdj", "CtJ0": "ctj0", "Cgs": "cgs", "CtJ1": "ctj1",
       "CT": "ct", "CtJ2": "ctj2", "CF": "cf", "Cpr[0]": "cpair0", "Cps[0]": "cspair0", "Cpr[r]": "cpairr",
       "Cps[r]": "cspairr"}


@pytest.mark.parametrize("nucleus", ["S40", "Gd162"])
def test_printed_couplings(nucleus):
    prob = golden_problem(nucleus, 0)
    printed = golden_points(nucleus, 0)[0]["couplings_printed"]
    c = prob.couplings
    for key, attr in MAP.items():
        assert abs(getattr(c, attr) - printed[key]) <= 6e-10 * max(1.0, abs(printed[key])), key
    assert c.sigma == printed["(sigma)"]


def test_density_dependent_couplings_on_grid(s40_k0):
    c = s40_k0.couplings
    rho = s40_k0.basis.rho_total
    assert np.allclose(c.crho, c.cr0 + c.crr * rho ** c.sigma, rtol=0, atol=1e-12)
    assert np.allclose(c.cpair, c.cpair0 + c.cpairr * rho, rtol=0, atol=1e-12)


def test_interaction_none():
    from pynfam_amd.interaction import init_interaction
    c = init_interaction("none", np.ones(4), np.array([-1.0, -1.0]), 0.5, 0.16)
    assert c.skip_residual and not np.any(c.crho)
