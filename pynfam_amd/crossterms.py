"""Cross-terms of a pnFAM solution: the response dR of operator F contracted
with the qp-basis matrix of another operator G of the same K and parity,
S_FG = -(1/pi) Tr(G dR) (pnfam_solver.f90:190-203 with contract_bbm,
pnfam_type_bbm.f90:393-421: the stored quadrant signs are not applied, so the
contraction is the plain sum over the X and Y quadrants)."""
from __future__ import annotations

import numpy as np


def contract(x, g20, g02):
    """-(1/pi) (sum g20 X + sum g02 Y) for the amplitudes x = [ReX, ReY, ImX, ImY]."""
    n = g20.shape[0]
    x = np.asarray(x).reshape(-1, 4, n)
    re = x[:, 0] @ g20 + x[:, 1] @ g02
    im = x[:, 2] @ g20 + x[:, 3] @ g02
    return -(re + 1j * im) / np.pi


def crossterm_problems(prob, **build_kw):
    """Problems of the cross-term operators of `prob` (empty for F and GT),
    evaluated on the block structure of `prob`'s operator."""
    from .extfield import CROSSTERMS, external_field
    from .problem import Problem
    out = []
    for g in CROSSTERMS.get(prob.operator, ()):
        st, elem = external_field(prob.basis, g, prob.K, struct=prob.f_struct)
        out.append((f"{prob.operator}x{g}", Problem(basis=prob.basis, couplings=prob.couplings, f_struct=st,
                                                    f_elem=elem, bminus=prob.bminus, operator=g, K=prob.K,
                                                    quench=prob.quench)))
    return out
