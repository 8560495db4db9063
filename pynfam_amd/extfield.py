"""External-field operators in the single-particle basis.

Restates the one-body allowed operators of pnfam_extfield.f90:
* init_fam_mapping (:763-840): the block pairing rule <1|F|2> with
  2*Omega_1 - 2*Omega_2 = 2K and the parity condition;
* ext_field_operator (:110-622), 'F' and 'GT' branches (:236-269) with the
  one-body (two_body_current_mode = 0) density factor of 1.
"""
from __future__ import annotations

import numpy as np

from .blocks import BlockStruct

PARITY_EVEN = {"F": True, "GT": True}
RANK = {"F": 0, "GT": 1}


def fam_mapping(basis, K, parity_even):
    """Block structure of the operator matrix (first matching block per row)."""
    db, nl, ns, npar = basis.db, basis.nl, basis.ns, basis.npar
    st = basis.isstart
    pty_blocks = all(np.all(npar[st[i]:st[i] + db[i]] == npar[st[i]]) for i in range(len(db)))
    ir2c = -np.ones(len(db), dtype=np.int64)
    for i1 in range(len(db)):
        p1 = st[i1]
        for i2 in range(len(db)):
            p2 = st[i2]
            if (2 * nl[p1] + ns[p1] - 2 * nl[p2] - ns[p2] == 2 * K) and \
                    (not pty_blocks or ((npar[p1] == npar[p2]) == parity_even)):
                ir2c[i1] = i2
                break
    if np.all(ir2c < 0):
        raise ValueError("The K and parity are outside the model space")
    return BlockStruct.from_map(db, ir2c)


def external_field(basis, label, K):
    """Returns (BlockStruct, elem) of the single-particle operator f_pn."""
    label = label.upper()
    if label not in PARITY_EVEN:
        raise ValueError(f"operator {label!r} is not implemented (one-body F and GT are)")
    if label == "F" and K != 0:
        raise ValueError("F only has K=0")
    if label == "GT" and abs(K) > 1:
        raise ValueError("GT has |K| <= 1")
    st = fam_mapping(basis, K, PARITY_EVEN[label])
    elem = np.zeros(st.nelem)
    wf, nl, ns = basis.wf, basis.nl, basis.ns
    for r, c, off in st.pairs():
        d1, d2 = basis.db[r], basis.db[c]
        s1, s2 = basis.isstart[r], basis.isstart[c]
        ov = wf[:, s1:s1 + d1].T @ wf[:, s2:s2 + d2]          # <1|2> on the grid
        l1, l2 = nl[s1:s1 + d1][:, None], nl[s2:s2 + d2][None, :]
        x1, x2 = ns[s1:s1 + d1][:, None], ns[s2:s2 + d2][None, :]
        if label == "F":
            m = np.where((l1 == l2) & (x1 == x2), ov, 0.0)
        elif K == 0:
            m = np.where((l1 == l2) & (x1 == x2), x1 * ov, 0.0)
        else:
            m = np.where((l1 == l2) & (x1 == x2 + 2 * K), -K * np.sqrt(2.0) * ov, 0.0)
        elem[off:off + d1 * d2] = m.ravel(order="F")
    return st, elem
