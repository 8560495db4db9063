"""The omega-independent inputs of one pnFAM (operator, K) solve.

Bundles what pnfam_solve sets up before its iteration (pnfam_solver.f90:32-59:
setup_pnfam -> HFB solution and couplings, setup_extfield -> operator f):
the doubled/spin-sorted HFB basis, the coupling constants on the grid and the
single-particle external field.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .blocks import BlockStruct
from .extfield import external_field
from .hfb import PnfamBasis, load_pnfam_basis
from .interaction import Couplings, init_interaction


@dataclass
class Problem:
    basis: PnfamBasis
    couplings: Couplings
    f_struct: BlockStruct
    f_elem: np.ndarray
    bminus: bool
    operator: str
    K: int
    quench: float = 1.0

    @property
    def nxy(self):
        return int(self.f_elem.shape[0])


def interaction_from_basis(basis: PnfamBasis, name="SKOP", **kw) -> Couplings:
    h = basis.hel
    return init_interaction(name, basis.rho_total, h.cpv0, float(h.cpv1[0]), float(h.nm["RHO_NM"]), **kw)


def build_problem(hel_path_or_basis, operator="GT", K=0, beta_type="-", interaction="SKOP",
                  quench_residual_int=1.0, energy_shift_prot=0.0, energy_shift_neut=0.0,
                  **interaction_kw) -> Problem:
    basis = hel_path_or_basis if isinstance(hel_path_or_basis, PnfamBasis) else load_pnfam_basis(hel_path_or_basis)
    if energy_shift_prot or energy_shift_neut:
        basis.Ep = np.where(basis.Ep != 0, basis.Ep, 0.0) + energy_shift_prot
        basis.En = np.where(basis.En != 0, basis.En, 0.0) + energy_shift_neut
    cpl = interaction_from_basis(basis, interaction, **interaction_kw)
    st, elem = external_field(basis, operator, K)
    quench = 0.0 if cpl.skip_residual else float(quench_residual_int)
    if beta_type not in ("-", "+"):
        raise ValueError(f"beta_type must be '-' or '+', got {beta_type!r}")
    return Problem(basis=basis, couplings=cpl, f_struct=st, f_elem=elem, bminus=beta_type == "-",
                   operator=operator.upper(), K=K, quench=quench)
