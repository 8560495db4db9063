import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def golden_points(nucleus, K):
    with open(os.path.join(GOLDEN, nucleus, f"GT-K{K}.json")) as fh:
        return json.load(fh)


_problems = {}


def golden_problem(nucleus, K, interaction=None):
    """Problem for an archived GT- set, with the namelist overrides of its .in files."""
    from pynfam_amd.problem import build_problem
    key = (nucleus, K, interaction)
    if key not in _problems:
        nml = golden_points(nucleus, K)[0]["namelist"]
        _problems[key] = build_problem(
            os.path.join(GOLDEN, nucleus, "hfbtho_output.hel"), "GT", K, nml["beta_type"],
            interaction or nml["interaction_name"], vpair_t0=nml["vpair_t0"], override_cs0=nml["override_cs0"],
            override_csr=nml["override_csr"], override_cds=nml["override_cds"])
    return _problems[key]


@pytest.fixture(scope="session")
def s40_k0():
    return golden_problem("S40", 0)


@pytest.fixture(scope="session")
def s40_k1():
    return golden_problem("S40", 1)
