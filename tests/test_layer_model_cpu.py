"""The torch layer model of the differentiable synthesis (adjoint.LayerModel,
run here on the CPU) against the host bundle producer synthetic.build_bundle:
populations per ion, Doppler widths, mass density, HCKT, H I population and
the continuum coefficients, at random abundance offsets."""

from __future__ import annotations

import numpy as np
import pytest


@pytest.mark.parametrize("teff,logg,mh", [(5750.0, 4.44, 0.0), (4000.0, 5.0, -1.0), (20000.0, 3.0, 0.3)])
def test_layer_model_matches_bundle(teff, logg, mh):
    import torch

    from kurucz_py_joss_submission_amd import synthetic as S
    from kurucz_py_joss_submission_amd.adjoint import LayerModel

    cols = S.make_atm(teff, logg, 72, mh)
    rng = np.random.default_rng(int(teff))
    delta = rng.normal(0.0, 0.3, len(S.ABUNDANCE_Z))
    ref = S.build_bundle(cols, dict(zip(S.ABUNDANCE_Z, delta.tolist())))
    lm = LayerModel(cols, device="cpu")
    T = torch.tensor(np.asarray(cols.temperature, np.float64))
    out = {k: v.numpy() for k, v in lm(T, torch.tensor(delta)).items()}

    def rel(a, b):
        return np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300))

    assert rel(out["pop"], ref.population_per_ion) < 1e-12
    assert rel(out["dop"], ref.doppler_per_ion) < 1e-13
    assert np.array_equal(out["pop"] == 0.0, ref.population_per_ion == 0.0)
    assert rel(out["rho"], ref.mass_density) < 1e-13
    assert rel(out["hckt"], ref.hckt) < 1e-14
    assert np.max(np.abs(out["cabs"] - ref.continuum_abs_coeff)) < 1e-10
    assert np.max(np.abs(out["cscat"] - ref.continuum_scat_coeff)) < 1e-10
    # TXNXN from the bundle's H I and He I populations
    txnxn = (ref.xnf_h + 0.42 * ref.xnf_he1) * (np.maximum(ref.temperature, 1.0) / 1e4) ** 0.3
    assert rel(out["txnxn"], txnxn) < 1e-12
