"""External-field operators in the single-particle basis.

Restates the one-body operators of pnfam_extfield.f90:
* init_fam_mapping (:763-840): the block pairing rule <1|F|2> with
  2*Omega_1 - 2*Omega_2 = 2K and the parity condition;
* init_external_field (:37-108): parities and ranks of the operators;
* ext_field_operator (:110-622), one-body branches (two_body_current_mode = 0):
  'F', 'GT' (:236-269) and the first-forbidden 'R', 'P', 'RS0', 'RS1', 'RS2',
  'PS0' (:319-620), as overlaps on the quadrature grid of the wavefunctions with
  z, r = 1/y, y, d/dz and d/dr;
* the cross-term operators of each first-forbidden operator
  (number_crossterms / setup_crossterms, :857-949): R, P and RS1 cross with
  R, RS1, P; RS0 and PS0 with RS0, PS0.
"""
from __future__ import annotations

import numpy as np

from .blocks import BlockStruct

PARITY_EVEN = {"F": True, "GT": True, "R": False, "P": False, "RS0": False, "RS1": False, "RS2": False,
               "PS0": False}
RANK = {"F": 0, "GT": 1, "R": 1, "P": 1, "RS0": 0, "RS1": 1, "RS2": 2, "PS0": 0}
MAX_K = {"F": 0, "GT": 1, "R": 1, "P": 1, "RS0": 0, "RS1": 1, "RS2": 2, "PS0": 0}
CROSSTERMS = {"R": ("R", "RS1", "P"), "P": ("R", "RS1", "P"), "RS1": ("R", "RS1", "P"),
              "RS0": ("RS0", "PS0"), "PS0": ("RS0", "PS0")}


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


def _block(label, K, ov, l1, l2, x1, x2):
    """One block of the operator from the grid overlaps ov[name] (d1 x d2) and the
    quantum numbers Lambda (l) and 2 Sigma (x) of the row / column states."""
    z, r, y, dz, dr, one = ov["z"], ov["r"], ov["y"], ov["dz"], ov["dr"], ov["1"]
    m = np.zeros_like(one)
    s2, s32, s3 = np.sqrt(2.0), np.sqrt(1.5), np.sqrt(3.0)
    same = (l1 == l2) & (x1 == x2)

    def put(mask, val):
        np.copyto(m, np.broadcast_to(val, m.shape), where=np.broadcast_to(mask, m.shape))

    if label == "F":
        put(same, one)
    elif label == "GT":
        if K == 0:
            put(same, x1 * one)
        else:
            put((l1 == l2) & (x1 == x2 + 2 * K), -K * s2 * one)
    elif label == "R":
        if K == 0:
            put(same, z)
        else:
            put((l1 == l2 + K) & (x1 == x2), -K / s2 * r)
    elif label == "P":
        if K == 0:
            put(same, -dz)
        else:
            put((l1 == l2 + K) & (x1 == x2), K / s2 * (dr - K * l2 * y))
    elif label == "RS0":
        put(same, -x1 * z)
        put((l1 == l2 + 1) & (x1 == x2 - 2), -r)
        put((l1 == l2 - 1) & (x1 == x2 + 2), -r)
    elif label == "RS1":
        if K == 0:
            put((l1 == l2 - 1) & (x1 == x2 + 2), s32 * r)
            put((l1 == l2 + 1) & (x1 == x2 - 2), -s32 * r)
        else:
            put((l1 == l2) & (x1 == x2 + 2 * K), s3 * z)
            put((l1 == l2 + K) & (x1 == x2), -x1 * s3 / 2.0 * r)
    elif label == "RS2":
        if K == 0:
            put(same, x1 * s2 * z)
            put((l1 == l2 + 1) & (x1 == x2 - 2), -1.0 / s2 * r)
            put((l1 == l2 - 1) & (x1 == x2 + 2), -1.0 / s2 * r)
        elif abs(K) == 1:
            put((l1 == l2) & (x1 == x2 + 2 * K), -K * s3 * z)
            put((l1 == l2 + K) & (x1 == x2), -K * x1 * s3 / 2.0 * r)
        else:
            put((l1 == l2 + K // 2) & (x1 == x2 + K), s3 * r)
    elif label == "PS0":
        put(same, -x1 * dz)
        put((l1 == l2 - 1) & (x1 == x2 + 2), -dr - l2 * y)
        put((l1 == l2 + 1) & (x1 == x2 - 2), -dr + l2 * y)
    return m


def external_field(basis, label, K, struct=None):
    """Returns (BlockStruct, elem) of the single-particle operator f_pn.  With
    `struct` given (a cross-term operator of the same K and parity) the
    operator is evaluated on that structure."""
    label = label.upper()
    if label not in PARITY_EVEN:
        raise ValueError(f"operator {label!r} is not implemented")
    if abs(K) > MAX_K[label]:
        raise ValueError(f"{label} has |K| <= {MAX_K[label]}")
    st = struct if struct is not None else fam_mapping(basis, K, PARITY_EVEN[label])
    elem = np.zeros(st.nelem)
    wf, nl, ns = basis.wf, basis.nl, basis.ns
    y, z = basis.grid.y, basis.grid.z
    for rb, cb, off in st.pairs():
        d1, d2 = basis.db[rb], basis.db[cb]
        s1, s2 = basis.isstart[rb], basis.isstart[cb]
        W1 = wf[:, s1:s1 + d1]
        W2 = wf[:, s2:s2 + d2]
        ov = {"1": W1.T @ W2}
        if label != "F" and label != "GT":
            ov.update({"z": W1.T @ (z[:, None] * W2), "r": W1.T @ (W2 / y[:, None]), "y": W1.T @ (y[:, None] * W2),
                       "dz": W1.T @ basis.wfdz[:, s2:s2 + d2], "dr": W1.T @ basis.wfdr[:, s2:s2 + d2]})
        else:
            zero = np.zeros_like(ov["1"])
            ov.update({"z": zero, "r": zero, "y": zero, "dz": zero, "dr": zero})
        l1, l2 = nl[s1:s1 + d1][:, None], nl[s2:s2 + d2][None, :]
        x1, x2 = ns[s1:s1 + d1][:, None], ns[s2:s2 + d2][None, :]
        m = _block(label, K, ov, l1, l2, x1, x2)
        elem[off:off + d1 * d2] = m.ravel(order="F")
    return st, elem
