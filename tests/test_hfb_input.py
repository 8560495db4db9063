"""HFB bridge: .hel reading and the rebuilt quasiparticle solution against the
reference's own HFBTHO listing (thoout.dat) for the archived S40 and Gd162 runs."""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN, golden_points
from pynfam_amd import hfb


@pytest.mark.parametrize("nucleus", ["S40", "Gd162"])
def test_fermi_levels_and_qp_energies(nucleus):
    b = hfb.load_pnfam_basis(os.path.join(GOLDEN, nucleus, "hfbtho_output.hel"))
    with open(os.path.join(GOLDEN, nucleus, "thoout_qp.json")) as fh:
        ref = json.load(fh)
    for it, lab in enumerate(["neutrons", "protons"]):
        sol = b.sols[it]
        # Fermi energy printed with 17 significant digits (hfbtho_solver.f90 resu)
        assert abs(sol.alast - ref[lab]["alast"]) < 1e-11 * abs(ref[lab]["alast"]) + 1e-12
        # all quasiparticle energies, printed with 6 decimals
        ours = np.sort(sol.E)
        theirs = np.sort([e[1] for e in ref[lab]["eqp_block_e"]])
        assert ours.shape == theirs.shape
        assert np.max(np.abs(ours - theirs)) <= 5.01e-7


@pytest.mark.parametrize("nucleus", ["S40", "Gd162"])
def test_basis_sizes_match_log(nucleus):
    b = hfb.load_pnfam_basis(os.path.join(GOLDEN, nucleus, "hfbtho_output.hel"))
    pt = golden_points(nucleus, 0)[0]
    assert b.dqp == pt["basis_size"]
    assert b.nb == pt["n_blocks"]
    assert int((b.db ** 2).sum()) == pt["nxy"]      # "Non-trivial HFB matrix elements" for K=0


def test_bogoliubov_unitarity_and_density_normalisation():
    b = hfb.load_pnfam_basis(os.path.join(GOLDEN, "S40", "hfbtho_output.hel"))
    h = b.hel
    # W = [U V; V U] per HFBTHO block is orthogonal: U^T U + V^T V = 1 on active columns
    for sol in b.sols:
        off = 0
        for ib in range(h.nb):
            nd = int(h.ID[ib])
            U, V = sol.U[ib], sol.V[ib]
            assert np.allclose(U.T @ U + V.T @ V, np.eye(nd), atol=1e-12)
            off += nd
    # particle numbers from the rebuilt density
    w = 1.0 / b.grid.wdcori
    assert abs(np.sum(b.sols[0].rho * w) - h.tz[0]) < 1e-10
    assert abs(np.sum(b.sols[1].rho * w) - h.tz[1]) < 1e-10


def test_wavefunctions_orthonormal_on_grid():
    b = hfb.load_pnfam_basis(os.path.join(GOLDEN, "S40", "hfbtho_output.hel"))
    for ib in range(b.nb):
        s, d = b.isstart[ib], b.db[ib]
        ov = b.wf[:, s:s + d].T @ b.wf[:, s:s + d]
        # states of one block with equal spin are orthonormal (quadrature built in)
        ns = b.ns[s:s + d]
        same = ns[:, None] == ns[None, :]
        assert np.allclose(np.where(same, ov, 0.0), np.eye(d), atol=1e-10)


def test_corrupt_hel_is_rejected(tmp_path):
    p = tmp_path / "bad.hel"
    raw = open(os.path.join(GOLDEN, "S40", "hfbtho_output.hel"), "rb").read()
    p.write_bytes(raw[:1000])
    with pytest.raises(hfb.HFBInputError):
        hfb.read_hel(str(p))
