"""Block-sparse matrix layout shared by every pnFAM matrix.

Same storage contract as the reference's `type(blockmatrix)`
(pnfam_type_blockmatrix.f90:19-26): at most one non-zero block per block row
and per block column; the non-zero blocks are stored one after another in
increasing block-row order, each block column-major.  Here the maps are
0-based and -1 marks "no block" (the reference uses 1-based maps and 0).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class BlockStruct:
    db: np.ndarray        # block sizes (nb,)
    ir2c: np.ndarray      # block row -> block column (-1: empty row)
    ir2m: np.ndarray      # block row -> offset of its block in elem (-1: empty)

    @property
    def nb(self):
        return len(self.db)

    @property
    def ic2r(self):
        out = -np.ones(self.nb, dtype=np.int64)
        for r, c in enumerate(self.ir2c):
            if c >= 0:
                out[c] = r
        return out

    @property
    def ic2m(self):
        out = -np.ones(self.nb, dtype=np.int64)
        for r, c in enumerate(self.ir2c):
            if c >= 0:
                out[c] = self.ir2m[r]
        return out

    @property
    def nelem(self):
        return int(sum(self.db[r] * self.db[c] for r, c in enumerate(self.ir2c) if c >= 0))

    def pairs(self):
        """(row block, col block, offset) for every stored block, in storage order."""
        return [(r, int(c), int(self.ir2m[r])) for r, c in enumerate(self.ir2c) if c >= 0]

    @staticmethod
    def from_map(db, ir2c):
        db = np.asarray(db, dtype=np.int64)
        ir2c = np.asarray(ir2c, dtype=np.int64)
        ir2m = -np.ones(len(db), dtype=np.int64)
        off = 0
        for r, c in enumerate(ir2c):
            if c >= 0:
                ir2m[r] = off
                off += db[r] * db[c]
        return BlockStruct(db=db, ir2c=ir2c, ir2m=ir2m)

    def transposed(self):
        return BlockStruct.from_map(self.db, self.ic2r)

    def to_dense(self, elem, isstart):
        n = int(self.db.sum())
        out = np.zeros((n, n), dtype=elem.dtype)
        for r, c, off in self.pairs():
            dr, dc = self.db[r], self.db[c]
            out[isstart[r]:isstart[r] + dr, isstart[c]:isstart[c] + dc] = \
                elem[off:off + dr * dc].reshape((dr, dc), order="F")
        return out

    def from_dense(self, dense, isstart):
        out = np.zeros(self.nelem, dtype=dense.dtype)
        for r, c, off in self.pairs():
            dr, dc = self.db[r], self.db[c]
            out[off:off + dr * dc] = dense[isstart[r]:isstart[r] + dr,
                                           isstart[c]:isstart[c] + dc].ravel(order="F")
        return out


def diagonal_struct(db):
    return BlockStruct.from_map(db, np.arange(len(db)))


def tr_struct(db):
    """Block i <-> its time-reversed partner i +- nb/2 (the V matrices,
    pnfam_setup.f90:314-320)."""
    nb = len(db)
    h = nb // 2
    return BlockStruct.from_map(db, np.concatenate([np.arange(h, nb), np.arange(0, h)]))


def product_struct(*structs):
    """Block map of A.B.C... (each factor untransposed)."""
    db = structs[0].db
    ir2c = []
    for r in range(len(db)):
        c = r
        for s in structs:
            c = s.ir2c[c] if c >= 0 else -1
        ir2c.append(c)
    return BlockStruct.from_map(db, ir2c)
