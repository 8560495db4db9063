"""Structurally exact, seeded synthetic HFB solutions -- for timing only.

No HFBTHO solution exists for the benchmark nuclei (Z=60 N=92 at N_sh=12/16,
Z=70 N=120 at N_sh=20), so throughput is measured on a stand-in with the exact
block structure of such a solution and physically sized values; it is never
used for parity.

* basis: the NOSCIL = (n+1)(n+2)(n+3)/6 lowest deformed HO states among
  N <= 50 by e = hw_z (n_z + 1/2) + hw_perp (2 n_r + Lambda + 1), with
  hw_z/hw_perp = q^-1, q = exp(3 sqrt(5/16 pi) beta), ties at the cut kept,
  blocks by (Omega, parity) -- the rule of hfbtho_solver.f90:999-1020, 3099-3160;
* grid: positive half of a 2*ngh-point Gauss-Hermite rule with doubled weights
  times e^{x^2}, ngl-point Gauss-Laguerre with weights times e^x (the
  quadrature the reference stores in .hel files);
* wavefunction tables: the same HO recurrences as real solutions (hfb.py);
* U, V: seeded random block-orthogonal rotations with BCS u, v;
  E = sqrt((e - lambda)^2 + Delta^2), zeroed above a 60 MeV pairing window;
* Fermi energies lambda_n = -2.0, lambda_p = -8.0 MeV (gaps 1.2, 1.0 MeV): a
  beta-minus window EQRPA_max = Delta M_nH + lambda_n - lambda_p = 6.78 MeV
  above the lowest two-quasiparticle energies, as for a neutron-rich emitter.
"""
from __future__ import annotations

import math

import numpy as np

from . import hfb

SEED = 20261202
LAMBDA_N, LAMBDA_P = -2.0, -8.0


def hfbtho_blocks(n00, beta, n00max=50):
    """Per (Omega, parity) block the list of (nr, nz, nl, ns) states, Omega>0."""
    q = math.exp(3.0 * math.sqrt(5.0 / (16.0 * math.pi)) * beta)
    hbz, hbp = q ** (-2.0 / 3.0), q ** (1.0 / 3.0)
    st = []
    for nz in range(n00max + 1):
        for nr in range(n00max // 2 + 1):
            for nl in range(n00max + 1):
                if nz + 2 * nr + nl > n00max:
                    continue
                for ns in (1, -1):
                    om2 = 2 * nl + ns
                    if om2 <= 0:
                        continue
                    e = hbz * (nz + 0.5) + hbp * (2 * nr + nl + 1)
                    st.append((e, nr, nz, nl, ns))
    noscil = (n00 + 1) * (n00 + 2) * (n00 + 3) // 6
    cut = sorted(s[0] for s in st)[noscil - 1] + 1e-9
    keep = sorted([s for s in st if s[0] < cut], key=lambda s: (s[0], s[1], s[2], s[3], s[4]))
    blocks = {}
    for e, nr, nz, nl, ns in keep:
        key = ((2 * nl + ns + 1) // 2, (nz + nl) % 2)
        blocks.setdefault(key, []).append((e, nr, nz, nl, ns))
    return [blocks[k] for k in sorted(blocks)]


def synthetic_hel(Z, N, n00, beta, ngh, ngl, seed=SEED):
    blocks = hfbtho_blocks(n00, beta)
    ID = np.array([len(b) for b in blocks], dtype=np.int64)
    qn = np.array([[s[1], s[2], s[3], s[4]] for b in blocks for s in b], dtype=np.int64)
    A = Z + N
    b0 = 1.005 * A ** (1.0 / 6.0)
    q = math.exp(3.0 * math.sqrt(5.0 / (16.0 * math.pi)) * beta)
    bp, bz = b0 * q ** (-1.0 / 6.0), b0 * q ** (1.0 / 3.0)
    x, w = np.polynomial.hermite.hermgauss(2 * ngh)
    xh, wh = x[x > 0], 2.0 * w[x > 0] * np.exp(x[x > 0] ** 2)
    xl, wl = np.polynomial.laguerre.laggauss(ngl)
    wl = wl * np.exp(xl)
    return hfb.HelData(Z=Z, N=N, skyrme="SKOP", pwi=60.0, nm={"RHO_NM": 0.1601687516},
                       crho=np.zeros(2), cdrho=np.zeros(2), ctau=np.zeros(2), crdr=np.zeros(2), crdj=np.zeros(2),
                       cj=np.zeros(2), cpv0=np.array([-253.75, -274.68]), cpv1=np.array([0.5, 0.5]), sigma=0.25,
                       hbzero=20.7355, b0=b0, bz=bz, bp=bp, n00=n00, nb=len(blocks), nt=int(ID.sum()),
                       ngh=ngh, ngl=ngl, xh=xh, xl=xl, wh=wh, wl=wl, ID=ID, qn=qn, si=0.0,
                       ala=np.array([-7.0, -8.0]), tz=np.array([float(N), float(Z)]), fields={}), blocks


def synthetic_basis(Z=60, N=92, n00=16, beta=0.3, ngh=40, ngl=40, seed=SEED) -> hfb.PnfamBasis:
    h, blocks = synthetic_hel(Z, N, n00, beta, ngh, ngl, seed)
    grid, qhla, fi1r, fi1z, fi2d = hfb.ho_wavefunctions(h)
    rng = np.random.default_rng(seed)
    hw0 = 41.0 * (Z + N) ** (-1.0 / 3.0)
    sols = []
    ia = np.concatenate([[0], np.cumsum(h.ID)[:-1]])
    for it, (lam, delta) in enumerate([(LAMBDA_N, 1.2), (LAMBDA_P, 1.0)]):
        E = np.zeros(h.nt)
        act = np.zeros(h.nt, dtype=bool)
        Us, Vs = [], []
        for ib, blk in enumerate(blocks):
            nd = len(blk)
            eps = np.sort(np.array([s[0] for s in blk])) * hw0 - 45.0   # HO single-particle spectrum
            O, _ = np.linalg.qr(rng.standard_normal((nd, nd)))
            Ek = np.sqrt((eps - lam) ** 2 + delta ** 2)
            v2 = 0.5 * (1.0 - (eps - lam) / Ek)
            u, v = np.sqrt(1.0 - v2), np.sqrt(v2)
            Us.append(O * u[None, :])
            Vs.append(O * v[None, :])
            E[ia[ib]:ia[ib] + nd] = Ek
            act[ia[ib]:ia[ib] + nd] = (eps <= h.pwi) | (Ek < h.pwi)
        # normal density from the active qp (as DENSIT's ro)
        tro = np.zeros(grid.nghl)
        ns = h.qn[:, 3]
        for ib in range(h.nb):
            nd = int(h.ID[ib])
            sl = slice(ia[ib], ia[ib] + nd)
            a_ = act[sl]
            Vb = Vs[ib][:, a_]
            up = ns[sl] > 0
            TU = qhla[:, sl][:, up] @ Vb[up]
            TD = qhla[:, sl][:, ~up] @ Vb[~up]
            tro += np.sum(TU * TU + TD * TD, axis=1)
        s = 2.0 * np.sum(tro)
        rho = tro * grid.wdcori * (2.0 * h.tz[it] / s)
        sols.append(hfb.QPSolution(E=E, U=Us, V=Vs, active=act, ala=lam, alast=lam, rho=rho))
    return hfb.assemble_pnfam_basis(h, grid, qhla, fi1r, fi1z, fi2d, sols)
