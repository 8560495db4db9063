"""HFB ground-state input for the pnFAM iteration: reads an HFBTHO `.hel` restart
file and rebuilds the quasiparticle solution (U, V, E) and the basis
wavefunction tables on the Gauss-Hermite x Gauss-Laguerre grid.

This is set-up work (once per nucleus), done on the host in FP64.  It follows
what the reference does when pnFAM starts (hfbtho_interface.f90:20-226 runs
HFBTHO for zero iterations from the restart file):

* `.hel` records: hfbtho_io.f90:487-738 (read_data, VERSION_DATA=5 keywords).
* HO wavefunctions with sqrt(quadrature weights) built in:
  hfbtho_solver.f90:3463-3540 (gaupol), 1215-1240 (coordinateLST),
  4237-4314 (optHFBTHO).
* HFB matrix from the mean and pairing fields on the grid:
  hfbtho_solver.f90:5145-5296 (gamdel).
* Block diagonalisation, Fermi level and pairing window:
  hfbtho_solver.f90:1441-1972 (hfbdiag), 1977-2075 (ALambda).
* Normal density on the grid for the density-dependent couplings:
  hfbtho_solver.f90:4318-4716 (DENSIT, the `ro` part only).
* pnFAM's doubling with explicit time-reversed states, spin sorting within
  blocks and the pairing-window cut: hfbtho_solution.f90:137-459.

Only the even-even, zero-temperature, no-constraint, HO-basis (not THO),
no-Lipkin-Nogami, zero-range (no Gogny / Gaussian Coulomb) case is supported;
anything else raises `HFBInputError`.
"""
from __future__ import annotations

import math
import struct
from dataclasses import dataclass, field

import numpy as np


class HFBInputError(RuntimeError):
    pass


# ---------------------------------------------------------------------------
# Fortran unformatted sequential records
# ---------------------------------------------------------------------------
def _records(path):
    with open(path, "rb") as fh:
        raw = fh.read()
    recs, p = [], 0
    while p < len(raw):
        if p + 4 > len(raw):
            raise HFBInputError(f"{path}: truncated record marker")
        (n,) = struct.unpack_from("<i", raw, p)
        if n < 0 or p + 8 + n > len(raw):
            raise HFBInputError(f"{path}: corrupt record at byte {p}")
        body = raw[p + 4:p + 4 + n]
        (n2,) = struct.unpack_from("<i", raw, p + 4 + n)
        if n2 != n:
            raise HFBInputError(f"{path}: record markers disagree at byte {p}")
        recs.append(body)
        p += 8 + n
    return recs


def _d(b):
    return np.frombuffer(b, dtype="<f8").copy()


def _i(b):
    return np.frombuffer(b, dtype="<i4").copy()


@dataclass
class HelData:
    Z: int
    N: int
    skyrme: str
    pwi: float
    nm: dict                 # nuclear-matter properties (RHO_NM, ...)
    crho: np.ndarray         # HFBTHO couplings (0:1) each
    cdrho: np.ndarray
    ctau: np.ndarray
    crdr: np.ndarray
    crdj: np.ndarray
    cj: np.ndarray
    cpv0: np.ndarray         # pairing strengths (n, p)
    cpv1: np.ndarray         # pairing density-dependence (n, p)
    sigma: float
    hbzero: float
    b0: float
    bz: float
    bp: float
    n00: int
    nb: int
    nt: int
    ngh: int
    ngl: int
    xh: np.ndarray
    xl: np.ndarray
    wh: np.ndarray
    wl: np.ndarray
    ID: np.ndarray           # states per HFBTHO block
    qn: np.ndarray           # (nt, 4): nr, nz, nl, ns
    si: float
    ala: np.ndarray          # Fermi levels (n, p)
    tz: np.ndarray           # particle numbers (n, p)
    fields: dict             # 'n'/'p' -> dict of 11 arrays (nghl)
    flags: dict = field(default_factory=dict)


def read_hel(path) -> HelData:
    """Read the keyword records of a VERSION_DATA=5 `.hel` file (hfbtho_io.f90:487-738)."""
    recs = _records(path)
    if len(recs) < 2 or _i(recs[0])[0] != 5:
        raise HFBInputError(f"{path}: not a VERSION_DATA=5 .hel file")
    out = {}
    i = 1
    while i < len(recs):
        r = recs[i]
        key = r.decode("latin-1") if len(r) == 8 else None
        if key is None:
            raise HFBInputError(f"{path}: expected a keyword record at #{i}")
        key = key.strip()
        i += 1
        if key == "Metadata":
            z, n = _i(recs[i]); i += 1
            fl1 = _i(recs[i]); i += 1
            fl2 = _i(recs[i]); i += 1
            out["Z"], out["N"] = int(z), int(n)
            out["flags"] = {"collective_inertia": bool(fl1[0]), "fission_fragments": bool(fl1[1]),
                            "pairing_regularization": bool(fl1[2]), "localization": bool(fl1[3]),
                            "switch_to_THO": int(fl2[0]), "projection": int(fl2[1]),
                            "temperature": bool(fl2[2]), "neck": bool(fl2[3])}
        elif key == "SkyFunct":
            b = recs[i]; i += 1
            logi = np.frombuffer(b[:20], dtype="<i4")
            out["skyrme"] = b[20:50].decode("latin-1").strip()
            out["finite_range"] = bool(logi[4])
            (out["pwi"],) = struct.unpack_from("<d", b, 54)
            nm = _d(recs[i]); i += 1
            out["nm"] = dict(zip(["E_NM", "K_NM", "SMASS_NM", "RHO_NM", "ASS_NM", "LASS_NM",
                                  "VMASS_NM", "P_NM", "KA_NM"], nm))
            c = _d(recs[i]); i += 1
            for k, name in enumerate(["crho", "cdrho", "ctau", "crdr", "crdj", "cj", "cpv0", "cpv1"]):
                out[name] = c[2 * k:2 * k + 2]
            out["sigma"] = c[16]
            hb = _d(recs[i]); i += 1
            out["hbzero"] = hb[0]
        elif key == "HO-Basis":
            out["b0"], out["bz"], out["bp"] = _d(recs[i]); i += 1
            n00, nb, nt, ngh, ngl, _nleg = _i(recs[i]); i += 1
            out.update(n00=int(n00), nb=int(nb), nt=int(nt), ngh=int(ngh), ngl=int(ngl))
            g = _d(recs[i]); i += 1
            out["xh"], out["xl"] = g[:ngh], g[ngh:ngh + ngl]
            out["wh"], out["wl"] = g[ngh + ngl:2 * ngh + ngl], g[2 * ngh + ngl:]
        elif key == "QuantNum":
            ID = _i(recs[i]); i += 1
            qn = np.array([_i(recs[i + k]) for k in range(int(ID.sum()))])
            i += int(ID.sum())
            out["ID"], out["qn"] = ID, qn
        elif key == "Various.":
            v1 = _d(recs[i]); v2 = _d(recs[i + 1]); v3 = _d(recs[i + 2]); i += 3
            out["si"] = v1[0]
            out["ala"] = v2[6:8]
            out["tz"] = v3[0:2]
            out["pwi_various"] = v2[0]
        elif key == "Constrai":
            ncons, _lmax = _i(recs[i]); i += 1
            if ncons != 0:
                raise HFBInputError("constrained HFB solutions are not supported")
            i += 4                      # multLambda, requested, qmoment, multLag
            if out.get("flags", {}).get("neck"):
                i += 3
            i += 1                      # qfield
        elif key == "Densits.":
            i += 2
        elif key in ("FieldsN.", "FieldsP."):
            names = ["v", "vhb", "vr", "vz", "vd", "vs", "vSFIZ", "vSZFI", "vSFIR", "vSRFI", "dv"]
            out.setdefault("fields", {})["n" if key == "FieldsN." else "p"] = {
                nm_: _d(recs[i + k]) for k, nm_ in enumerate(names)}
            i += len(names)
        elif key == "Blocking":
            i += 2
        elif key == "Blk-rest":
            i += 1
        elif key == "Temperat":
            raise HFBInputError("finite-temperature HFB solutions are not supported")
        elif key == "THObasis":
            raise HFBInputError("THO-basis HFB solutions are not supported")
        elif key == "HFBmatrX":
            raise HFBInputError("HFB-matrix records (Gogny) are not supported")
        else:
            raise HFBInputError(f"unsupported .hel keyword {key!r}")
    if out.get("finite_range"):
        raise HFBInputError("finite-range functionals are not supported")
    if out["flags"]["switch_to_THO"] != 0 or out["flags"]["temperature"]:
        raise HFBInputError("only HO-basis, zero-temperature solutions are supported")
    out.pop("finite_range", None)
    out.pop("pwi_various", None)
    return HelData(**out)


# ---------------------------------------------------------------------------
# HO basis wavefunctions on the quadrature grid
# ---------------------------------------------------------------------------
@dataclass
class Grid:
    ngh: int
    ngl: int
    nghl: int
    y: np.ndarray        # 1/r on the grid [fm^-1]
    z: np.ndarray        # z on the grid [fm]
    wdcori: np.ndarray   # inverse quadrature weight times Jacobian


def ho_wavefunctions(h: HelData):
    """HO basis functions, their r/z derivatives and the (r,z) Laplacian on the
    grid, each with sqrt(weights) mixed in (gaupol + optHFBTHO, HO branch).
    Returns (grid, qhla, fi1r, fi1z, fi2d) with arrays shaped (nghl, nt) and the
    grid index ihil = ih + il*ngh (ih fastest)."""
    ngh, ngl = h.ngh, h.ngl
    nr, nz, nl = h.qn[:, 0], h.qn[:, 1], h.qn[:, 2]
    nzm, nrm, nlm = int(nz.max()), int(nr.max()), int(nl.max())
    xh, xl, wh, wl = h.xh, h.xl, h.wh, h.wl
    sq = np.sqrt(np.arange(0, 2 * (nzm + nrm + nlm) + 8, dtype=np.float64))
    sqi = np.zeros_like(sq)
    sqi[1:] = 1.0 / sq[1:]
    wfi = np.ones(nlm + 3)
    for k in range(1, nlm + 3):
        wfi[k] = wfi[k - 1] / sq[k]
    # Hermite functions qh(n, ih) and derivatives qh1
    qh = np.zeros((max(nzm, 1) + 1, ngh))
    qh1 = np.zeros_like(qh)
    w4pii = math.pi ** (-0.25)
    w0 = w4pii * np.exp(-0.5 * xh * xh) * np.sqrt(wh)
    qh[0] = w0
    qh[1] = sq[2] * w0 * xh
    qh1[0] = -w0 * xh
    qh1[1] = sq[2] * w0 * (1.0 - xh * xh)
    for n in range(2, nzm + 1):
        qh[n] = sqi[n] * (sq[2] * xh * qh[n - 1] - sq[n - 1] * qh[n - 2])
        qh1[n] = sq[n + n] * qh[n - 1] - xh * qh[n]
    # Laguerre functions ql(n, l, il) and derivatives ql1
    ql = np.zeros((max(nrm, 1) + 1, nlm + 1, ngl))
    ql1 = np.zeros_like(ql)
    for l in range(nlm + 1):
        w00 = sq[2] * np.exp(-0.5 * xl)
        w0 = w00 * np.sqrt(0.5 * wl * xl ** l)
        ql[0, l] = wfi[l] * w0
        ql[1, l] = (l + 1 - xl) * wfi[l + 1] * w0
        ql1[0, l] = (l - xl) * wfi[l] * w0
        ql1[1, l] = (float(l * l + l) - xl * float(l + l + 3) + xl * xl) * wfi[l + 1] * w0
        for n in range(2, nrm + 1):
            dsq = sq[n] * sq[n + l]
            d1 = float(n + n + l - 1) - xl
            d2 = sq[n - 1] * sq[n - 1 + l]
            d3 = n + n + l - xl
            d4 = 2.0 * dsq
            ql[n, l] = (d1 * ql[n - 1, l] - d2 * ql[n - 2, l]) / dsq
            ql1[n, l] = d3 * ql[n, l] - d4 * ql[n - 1, l]
    bz, bp = h.bz, h.bp
    ihh = np.tile(np.arange(ngh), ngl)           # ih for each ihil
    ill = np.repeat(np.arange(ngl), ngh)         # il for each ihil
    XH = xh[ihh]
    XL = xl[ill]
    y = 1.0 / (np.sqrt(XL) * bp)
    v2 = 0.5 / XL
    v4 = v2 * v2
    QH = qh[nz][:, ihh]                           # (nt, nghl)
    QH1 = qh1[nz][:, ihh]
    QL = ql[nr, nl][:, ill]
    QL1 = ql1[nr, nl][:, ill]
    qhla = QH * QL
    qhl1a = QH * QL1 * v2
    qh1la = QH1 * QL
    fi1r = (2.0 * np.sqrt(XL) / bp) * qhl1a
    fi1z = qh1la / bz
    cnz = (2 * nz + 1)[:, None].astype(np.float64)
    cnr = (2 * nr + nl + 1)[:, None].astype(np.float64)
    sml2 = (nl * nl)[:, None].astype(np.float64)
    fi2d = ((XH * XH - cnz) / (bz * bz) + 4.0 * (0.25 - cnr * v2 + sml2 * v4) * XL / (bp * bp)) * qhla
    wdcor = math.pi * wh[ihh] * wl[ill] * bz * bp * bp
    grid = Grid(ngh=ngh, ngl=ngl, nghl=ngh * ngl, y=y, z=bz * XH, wdcori=1.0 / wdcor)
    return grid, qhla.T.copy(), fi1r.T.copy(), fi1z.T.copy(), fi2d.T.copy()


# ---------------------------------------------------------------------------
# HFB matrix from the fields (gamdel) and its block diagonalisation (hfbdiag)
# ---------------------------------------------------------------------------
def _block_omega_factors(nb):
    """LAPLUS per block for reflection-symmetric bases (hfbtho_solver.f90:5214-5219):
    blocks come in (Omega, parity) pairs, so LAPLUS=(ib+1)/2 with 1-based ib."""
    return [((ib + 1) + 1) // 2 for ib in range(nb)]


def hfb_matrices(h: HelData, grid: Grid, qhla, fi1r, fi1z, fi2d):
    """Per-block mean-field h and pairing Delta matrices for neutrons (it=0) and
    protons (it=1), following gamdel's matrix elements (lower triangle there;
    here the full symmetric matrices)."""
    ns = h.qn[:, 3]
    ia = np.concatenate([[0], np.cumsum(h.ID)[:-1]])
    lap = _block_omega_factors(h.nb)
    y = grid.y
    out = []
    for ib in range(h.nb):
        nd = int(h.ID[ib])
        sl = slice(ia[ib], ia[ib] + nd)
        xlap = float(lap[ib])
        xlam = xlap - 1.0
        xlamy, xlapy = xlam * y, xlap * y
        xlamy2, xlapy2 = xlam * xlam * y * y, xlap * xlap * y * y
        xlampy = xlamy + xlapy
        su = (ns[sl] > 0).astype(np.float64)[:, None]
        sd = (ns[sl] < 0).astype(np.float64)[:, None]
        Q, R, Zd, D2 = qhla[:, sl].T, fi1r[:, sl].T, fi1z[:, sl].T, fi2d[:, sl].T   # (nd, nghl)
        FU, FUR, FUZ = Q * su, R * su, Zd * su
        FD, FDR, FDZ = Q * sd, R * sd, Zd * sd
        FUD2 = (D2 - xlamy2 * Q) * su
        FDD2 = (D2 - xlapy2 * Q) * sd
        per_it = []
        for it in ("n", "p"):
            f = h.fields[it]
            v, vhb, vr, vz, vd, vs = f["v"], f["vhb"], f["vr"], f["vz"], f["vd"], f["vs"]
            vSFIZ, vSZFI, vSFIR, vSRFI, dv = f["vSFIZ"], f["vSZFI"], f["vSFIR"], f["vSRFI"], f["dv"]

            def bil(A, w, B):
                return (A * w) @ B.T

            # spin up-up
            hm = np.zeros((nd, nd))
            # vh*v + hbh*vhb + vdh*vd + SNABLAR*vr + SNABLAZ*vz + vsh*vs + SFIZ*vSFIZ
            hm += bil(FU, v + xlamy2 * vhb + 2.0 * xlamy2 * vd + 2.0 * xlamy * vSFIZ, FU)
            hm += bil(FUR, vhb + 2.0 * vd, FUR) + bil(FUZ, vhb + 2.0 * vd, FUZ)
            hm += bil(FU, vd, FUD2) + bil(FUD2, vd, FU)
            hm += bil(FUR, vr + xlamy * vs, FU) + bil(FU, vr + xlamy * vs, FUR)
            hm += bil(FUZ, vz, FU) + bil(FU, vz, FUZ)
            # spin down-down
            hm += bil(FD, v + xlapy2 * vhb + 2.0 * xlapy2 * vd - 2.0 * xlapy * vSFIZ, FD)
            hm += bil(FDR, vhb + 2.0 * vd, FDR) + bil(FDZ, vhb + 2.0 * vd, FDZ)
            hm += bil(FD, vd, FDD2) + bil(FDD2, vd, FD)
            hm += bil(FDR, vr - xlapy * vs, FD) + bil(FD, vr - xlapy * vs, FDR)
            hm += bil(FDZ, vz, FD) + bil(FD, vz, FDZ)
            # spin flip, element (a=1, b=2) with a down / b up ("DoUp" in gamdel)
            #  vsh  = -FDR1*FUZ2 + FUR2*FDZ1 - FDZ1*FU2*xlamy - FUZ2*FD1*xlapy
            #  SRFI = -FDR1*FU2 + FUR2*FD1 ;  SFIR = FD1*FU2*xlampy ;  SZFI = -FDZ1*FU2 + FUZ2*FD1
            du = (-bil(FDR, vs, FUZ) + bil(FDZ, vs, FUR) - bil(FDZ, vs * xlamy, FU) - bil(FD, vs * xlapy, FUZ)
                  - bil(FDR, vSRFI, FU) + bil(FD, vSRFI, FUR)
                  + bil(FD, vSFIR * xlampy, FU)
                  - bil(FDZ, vSZFI, FU) + bil(FD, vSZFI, FUZ))
            # element (a=1, b=2) with a up / b down ("UpDo"):
            #  vsh  = FUR1*FDZ2 - FDR2*FUZ1 - FUZ1*FD2*xlapy - FDZ2*FU1*xlamy
            #  SRFI = FUR1*FD2 - FDR2*FU1 ;  SFIR = FU1*FD2*xlampy ;  SZFI = -FDZ2*FU1 + FUZ1*FD2
            ud = (bil(FUR, vs, FDZ) - bil(FUZ, vs, FDR) - bil(FUZ, vs * xlapy, FD) - bil(FU, vs * xlamy, FDZ)
                  + bil(FUR, vSRFI, FD) - bil(FU, vSRFI, FDR)
                  + bil(FU, vSFIR * xlampy, FD)
                  - bil(FU, vSZFI, FDZ) + bil(FUZ, vSZFI, FD))
            # gamdel fills only the lower triangle (n1 >= n2, a=n1, b=n2); the
            # matrix is then used symmetrically (DSYEVR on the 'L' triangle).
            low = np.tril(np.ones((nd, nd), dtype=bool))
            flip = np.where(low, du + ud, 0.0)
            same = np.where(low, hm, 0.0)
            hl = same + flip
            hmat = hl + np.tril(hl, -1).T
            dl = np.where(low, bil(FU, dv, FU) + bil(FD, dv, FD), 0.0)
            dmat = dl + np.tril(dl, -1).T
            per_it.append((hmat, dmat))
        out.append(per_it)
    return out


CUTOFF_TOL = 1.0e-6


def _alambda(al, erhfb, drhfb, tz, cpv0_it):
    """Fermi level from the particle-number condition (ALambda, pairing branch,
    hfbtho_solver.f90:1977-2075, zero temperature, no blocking)."""
    if cpv0_it == 0.0:
        ntz = int(tz + 0.1) // 2
        d = np.sort(erhfb)
        return 0.5 * (d[ntz - 1] + d[ntz]) if ntz < len(d) else d[ntz - 1] + 0.001
    xinf, xsup = -1000.0, 1000.0
    for lit in range(1, 501):
        yv = erhfb - al
        a = yv * yv + drhfb ** 2
        b = np.sqrt(a)
        vh = np.where(b > 0, 0.5 * (1.0 - yv / np.where(b > 0, b, 1.0)), 0.0)
        vh = np.where(vh < 1.0e-12, 0.0, vh)
        vh = np.where(vh - 1.0 > 1.0e-12, 1.0, vh)
        dvh = np.where(b > 0, 0.5 * drhfb ** 2 / np.where(b > 0, a * b, 1.0), 0.0)
        sn = 0.0
        dez = 0.0
        for k in range(len(vh)):          # sequential sums, as in the reference loop
            sn += 2.0 * vh[k]
            dez += 2.0 * dvh[k]
        ez = sn - tz
        absez = abs(ez) / tz
        if ez < 0.0:
            xinf = max(xinf, al)
        else:
            xsup = min(xsup, al)
        if lit == 1:
            al = al - ez if absez <= 0.10 else al - 0.10 * math.copysign(1.0, ez)
        else:
            al = al - ez / (dez + 1.0e-20)
        if al < xinf or al > xsup:
            al = 0.5 * (xinf + xsup)
        if absez <= 1.0e-10:
            return al
    return al


@dataclass
class QPSolution:
    """Bogoliubov solution of one isospin in HFBTHO's block ordering."""
    E: np.ndarray            # (nt,)  qp energy per (block, k)
    U: list                  # per block (nd, nd): U[a, k]
    V: list
    active: np.ndarray       # (nt,) bool: below the pairing window
    ala: float
    alast: float
    rho: np.ndarray          # normal density on the grid (nghl,)


def hfb_diagonalise(h: HelData, grid: Grid, mats, qhla):
    ns = h.qn[:, 3]
    ia = np.concatenate([[0], np.cumsum(h.ID)[:-1]])
    sitest = min(0.10, h.si * 0.010)
    sols = []
    for it in range(2):
        tz = float(h.tz[it])
        al = float(h.ala[it])
        sumnz = 1.0
        inner = -1
        norm_to_improve = True
        while norm_to_improve:
            inner += 1
            if abs(sumnz) < sitest or inner == 20:
                norm_to_improve = False
            sumnz = 0.0
            erh, drh = [], []
            E = np.zeros(int(h.ID.sum()))
            act = np.zeros(int(h.ID.sum()), dtype=bool)
            Us, Vs = [], []
            for ib in range(h.nb):
                nd = int(h.ID[ib])
                hm, dm = mats[ib][it]
                big = np.zeros((2 * nd, 2 * nd))
                big[:nd, :nd] = hm - al * np.eye(nd)
                big[nd:, nd:] = -hm + al * np.eye(nd)
                big[nd:, :nd] = dm
                big[:nd, nd:] = dm
                w, vec = np.linalg.eigh(big)
                Us.append(vec[:nd, nd:].copy())
                Vs.append(vec[nd:, nd:].copy())
                for k in range(nd):
                    pn = float(np.sum(vec[nd:, nd + k] ** 2))
                    eqpe = float(w[nd + k])
                    ela = eqpe * (1.0 - 2.0 * pn)
                    enb = ela + al
                    ekb = math.sqrt(abs(eqpe ** 2 - ela ** 2))
                    arg = 100.0 * (enb - h.pwi)
                    expo = math.exp(arg) if abs(arg) < math.log(np.finfo(float).max) else np.inf
                    lpr = enb <= h.pwi or abs(1.0 / (1.0 + expo)) > CUTOFF_TOL
                    E[ia[ib] + k] = eqpe
                    if lpr:
                        act[ia[ib] + k] = True
                        erh.append(enb)
                        drh.append(ekb)
                        sumnz += 2.0 * pn
            sumnz -= tz
            final = (E.copy(), [u.copy() for u in Us], [v.copy() for v in Vs], act.copy(), al)
            al = _alambda(al, np.array(erh), np.array(drh), tz, float(h.cpv0[it]))
        E, Us, Vs, act, al_used = final
        # Normal density (DENSIT, ro only): sum over active qp of (sum_a V_ak phi_a)^2 per spin
        tro = np.zeros(grid.nghl)
        for ib in range(h.nb):
            nd = int(h.ID[ib])
            sl = slice(ia[ib], ia[ib] + nd)
            a_ = act[sl]
            if not a_.any():
                continue
            Vb = Vs[ib][:, a_]
            up = ns[sl] > 0
            TU = qhla[:, sl][:, up] @ Vb[up]
            TD = qhla[:, sl][:, ~up] @ Vb[~up]
            tro += np.sum(TU * TU + TD * TD, axis=1)
        s = 2.0 * np.sum(tro)
        rho = tro * grid.wdcori * (2.0 * tz / s)
        sols.append(QPSolution(E=E, U=Us, V=Vs, active=act, ala=al_used, alast=al, rho=rho))
    return sols


# ---------------------------------------------------------------------------
# pnFAM view of the solution: explicit time-reversed blocks, spin-sorted rows
# ---------------------------------------------------------------------------
@dataclass
class PnfamBasis:
    """Everything the pnFAM iteration reads from the HFB solution, in pnFAM's
    doubled (Omega>0 blocks then their time-reversed copies) and spin-sorted
    basis.  Arrays are FP64; matrices use the block-matrix storage of
    pnfam_type_blockmatrix.f90:19-26 (blocks in row order, column-major inside)."""
    nb: int
    db: np.ndarray            # (nb,) block sizes
    isstart: np.ndarray       # (nb,) 0-based first state of each block
    nr: np.ndarray
    nz: np.ndarray
    nl: np.ndarray
    ns: np.ndarray
    npar: np.ndarray
    num_spin_up: np.ndarray   # (nb,)
    grid: Grid
    wf: np.ndarray            # (nghl, dqp) Fortran-style column per state
    wfdr: np.ndarray
    wfdz: np.ndarray
    wfdp: np.ndarray          # y * nl * wf
    wfd2_all: np.ndarray      # (r,z) Laplacian minus y^2 nl^2 wf
    Ep: np.ndarray            # (dqp,) 0 above the pairing window
    En: np.ndarray
    Up: np.ndarray            # (sum db^2,) block-diagonal storage
    Vp: np.ndarray            # rows of block i -> columns of block i +- nb/2
    Un: np.ndarray
    Vn: np.ndarray
    rho_total: np.ndarray     # isoscalar normal density on the grid
    hel: HelData = None
    sols: list = None

    @property
    def dqp(self):
        return int(self.db.sum())

    @property
    def nghl(self):
        return self.grid.nghl


def load_pnfam_basis(hel_path) -> PnfamBasis:
    h = read_hel(hel_path)
    if h.N % 2 or h.Z % 2:
        raise HFBInputError("odd nuclei (blocking / EFA) are not supported")
    grid, qhla, fi1r, fi1z, fi2d = ho_wavefunctions(h)
    mats = hfb_matrices(h, grid, qhla, fi1r, fi1z, fi2d)
    sols = hfb_diagonalise(h, grid, mats, qhla)
    return assemble_pnfam_basis(h, grid, qhla, fi1r, fi1z, fi2d, sols)


def assemble_pnfam_basis(h, grid, qhla, fi1r, fi1z, fi2d, sols) -> PnfamBasis:
    """Doubling, spin sort and pairing-window cut (hfbtho_solution.f90:162-459)."""
    nbh = h.nb
    nb = 2 * nbh
    dh = h.ID.astype(np.int64)
    db = np.concatenate([dh, dh])
    nth = int(dh.sum())
    dqp = 2 * nth
    nr = np.concatenate([h.qn[:, 0], h.qn[:, 0]])
    nz = np.concatenate([h.qn[:, 1], h.qn[:, 1]])
    nl = np.concatenate([h.qn[:, 2], -h.qn[:, 2]])
    ns = np.concatenate([h.qn[:, 3], -h.qn[:, 3]])
    # parity of the HO state: (nz + nl) even -> 1, odd -> 2
    par = np.where((h.qn[:, 1] + h.qn[:, 2]) % 2 == 0, 1, 2)
    npar = np.concatenate([par, par])
    sgn = np.where(h.qn[:, 3] < 0, -1.0, 1.0)
    wf = np.concatenate([qhla, qhla * sgn], axis=1)
    wfdr = np.concatenate([fi1r, fi1r * sgn], axis=1)
    wfdz = np.concatenate([fi1z, fi1z * sgn], axis=1)
    wfd2 = np.concatenate([fi2d, fi2d * sgn], axis=1)

    def bogoliubov(sol):
        E = np.concatenate([sol.E, sol.E])
        Ub = [u for u in sol.U] + [u for u in sol.U]
        Vb = [-v for v in sol.V] + [v for v in sol.V]
        act = np.concatenate([sol.active, sol.active])
        return E, Ub, Vb, act

    En, Unb, Vnb, actn = bogoliubov(sols[0])
    Ep, Upb, Vpb, actp = bogoliubov(sols[1])
    isstart = np.concatenate([[0], np.cumsum(db)[:-1]])
    # spin sort: spin-up rows first inside every block (hfbtho_solution.f90:305-343)
    new_order = np.zeros(dqp, dtype=np.int64)
    nsu = np.zeros(nb, dtype=np.int64)
    for ib in range(nb):
        idx = np.arange(isstart[ib], isstart[ib] + db[ib])
        up = idx[ns[idx] > 0]
        dn = idx[ns[idx] < 0]
        nsu[ib] = len(up)
        new_order[isstart[ib]:isstart[ib] + db[ib]] = np.concatenate([up, dn])
    nr, nz, nl, ns, npar = nr[new_order], nz[new_order], nl[new_order], ns[new_order], npar[new_order]
    wf, wfdr, wfdz, wfd2 = wf[:, new_order], wfdr[:, new_order], wfdz[:, new_order], wfd2[:, new_order]

    def sort_rows(blocks):
        out = []
        for ib, B in enumerate(blocks):
            loc = new_order[isstart[ib]:isstart[ib] + db[ib]] - isstart[ib]
            out.append(B[loc, :])
        return out

    Unb, Vnb, Upb, Vpb = sort_rows(Unb), sort_rows(Vnb), sort_rows(Upb), sort_rows(Vpb)

    def erase(E, blocks, act):
        E = E.copy()
        E[~act] = 0.0
        res = []
        for ib, B in enumerate(blocks):
            a = act[isstart[ib]:isstart[ib] + db[ib]]
            B = B.copy()
            B[:, ~a] = 0.0
            res.append(B)
        return E, res

    # erase_above_pwi (hfbtho_solution.f90:402-459): activity of qp column k of
    # storage block ib is that of qp (ib, k); TR partners share activity.
    En, Unb = erase(En, Unb, actn)
    _, Vnb = erase(En, Vnb, actn)
    Ep, Upb = erase(Ep, Upb, actp)
    _, Vpb = erase(Ep, Vpb, actp)

    def flat(blocks):
        return np.concatenate([B.ravel(order="F") for B in blocks])

    y = grid.y
    wfdp = y[:, None] * nl[None, :] * wf
    wfd2_all = wfd2 - y[:, None] * nl[None, :] * wfdp
    return PnfamBasis(nb=nb, db=db, isstart=isstart, nr=nr, nz=nz, nl=nl, ns=ns, npar=npar,
                      num_spin_up=nsu, grid=grid,
                      wf=np.asfortranarray(wf), wfdr=np.asfortranarray(wfdr),
                      wfdz=np.asfortranarray(wfdz), wfdp=np.asfortranarray(wfdp),
                      wfd2_all=np.asfortranarray(wfd2_all),
                      Ep=Ep, En=En, Up=flat(Upb), Vp=flat(Vpb), Un=flat(Unb), Vn=flat(Vnb),
                      rho_total=sols[0].rho + sols[1].rho, hel=h, sols=sols)
