"""Full-window configuration at full size against the committed oracle fixture.

Teff 4000 K, log g 5.0, 72 layers, 1,000,000 lines, 300-1800 nm, R = 300000:
W = 537,529, kappa grid 72 x 537,529.  The fixture (tests/golden/
make_c2_fixture.py) holds the oracle's flux and continuum at every 97th
wavelength and the exact evaluation / iteration counts.
"""

from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

FIX = Path(__file__).resolve().parent / "golden" / "c2_oracle_sample.npz"


@pytest.mark.gpu
def test_c2_full_window_parity(gpu):
    if not FIX.exists():
        pytest.skip("fixture not generated")
    from kurucz_py_joss_submission_amd import synthetic
    from kurucz_py_joss_submission_amd.device import SpectrumPlan

    z = np.load(FIX)
    wk = synthetic.make_workload("c2", 4000.0, 5.0, 1_000_000, 300.0, 1800.0, seed=1)
    plan = SpectrumPlan(wk.atmosphere, wk.catalog, 300.0, 1800.0, 300000.0)
    plan.run(stats=True)
    assert plan.W == int(z["n_wavelengths"]) == 537529
    st = plan.stats.cpu().numpy()
    assert st[0] == z["e_line"] and st[1] == z["wing_pairs"]
    assert st[2] == z["sum_maxstep"] and st[3] == z["max_maxstep"]
    idx = z["index"]
    flux = plan.flux.cpu().numpy()
    cont = plan.cont.cpu().numpy()
    it = plan.iters.cpu().numpy()
    assert np.array_equal(plan.grid[idx], z["wavelength"])
    rel_c = np.max(np.abs(cont[idx] - z["continuum"]) / np.abs(z["continuum"]))
    floor = 1e-3 * np.median(z["continuum"])
    rel_f = np.max(np.abs(flux[idx] - z["flux"]) / np.maximum(np.abs(z["flux"]), floor))
    assert rel_c < 1e-6 and rel_f < 1e-6, (rel_c, rel_f)
    assert np.mean(it[idx] == z["iters"]) > 0.999
    assert abs(float(np.sum(flux)) / float(z["flux_sum"]) - 1) < 1e-9
