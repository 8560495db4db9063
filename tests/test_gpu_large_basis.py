"""Mixed path on a production-size basis (N_sh=16, 40x40 grid, dqp=1958): the
synthetic HFB input of the throughput workload, GT- K=0, 32 points of the CIRCLE
contour.  At this size the kernels take the paths the small archived bases never
reach: split-K density GEMMs (96 or more columns), one strip per wave in the
qp->sp transforms (blocks of four or more 16-row strips), the factored grid
stage for the last few points, the FP32 Broyden history over long tails.
Reference: the FP64 path on the same points.  Tolerance 1e-4 relative in S
(MIXED_NOREFINE), every point converged, mean iteration count within 20 % of
FP64 (single points near the real axis may take longer: the reduced-precision
increments slow their last digits)."""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def workload():
    sys.path.insert(0, ROOT)
    import bench
    from pynfam_amd import engine as E
    from pynfam_amd.rate import eqrpa_max_beta_minus
    basis, probs = bench.build_workload(16, 0.3, bench.Z_PARENT, bench.N_PARENT, 40)
    emax = eqrpa_max_beta_minus(basis.sols[0].ala, basis.sols[1].ala)
    w = bench.contour(128, emax)[::4]
    eng = E.GpuEngine(probs[0])
    ref = eng.solve(w, precision=E.FP64)
    eng.close()
    assert np.all(ref["iterations"] > 0), ref["iterations"]
    return probs[0], w, ref


@pytest.mark.parametrize("grid", ["auto", "tables"])
def test_mixed_large_basis_against_fp64(workload, grid, monkeypatch):
    from pynfam_amd import engine as E
    prob, w, ref = workload
    monkeypatch.setenv("PNFAM_GPU_MIX_GRID", grid)
    eng = E.GpuEngine(prob)
    r = eng.solve(w, precision=E.MIXED_NOREFINE)
    eng.close()
    assert np.all(r["iterations"] > 0), r["iterations"]
    err = np.abs(r["strength"] - ref["strength"]) / np.abs(ref["strength"])
    print(f"N_sh=16 K=0 {grid}: max rel err {err.max():.2e}, median {np.median(err):.2e}, "
          f"iterations {r['iterations'].tolist()} (FP64 {ref['iterations'].tolist()})")
    assert err.max() < 1e-4, err.max()
    assert r["iterations"].mean() <= 1.2 * ref["iterations"].mean(), (r["iterations"].mean(), ref["iterations"].mean())
