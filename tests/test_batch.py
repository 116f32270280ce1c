"""Batched spectra: uneven depth counts and windows in one plan (graph replay,
autotuned persistent gather) give the same bits as one SpectrumPlan per
instance, on every graph replay."""

from __future__ import annotations

import numpy as np
import pytest

INSTANCES = [  # Teff, log g, layers, [M/H], window (nm)
    (5750.0, 4.44, 72, 0.0, 500.0, 501.5),
    (4000.0, 5.0, 64, -1.0, 600.2, 601.0),
    (9000.0, 3.5, 80, 0.3, 450.0, 452.0),
    (20000.0, 2.0, 70, -2.0, 680.0, 680.7),
    (3500.0, 0.5, 77, 0.5, 420.0, 420.4),
]


class _Inputs(list):
    """(atmosphere, catalog, wl0, wl1) per instance, plus the oracle's catalog."""

    oracle_catalog = None


@pytest.fixture(scope="module")
def batch_inputs():
    import catalog as C
    from kurucz_py_joss_submission_amd import lines as L
    from kurucz_py_joss_submission_amd import synthetic as S

    import tempfile
    from pathlib import Path

    d = S.synthetic_lines(40_000, 400.0, 700.0, seed=5)
    with tempfile.TemporaryDirectory() as td:
        gf = Path(td) / "lines.gfall"
        S.write_line_list(gf, d)
        cat = L.load_catalog(gf)
        ocat = C.read_gfall(gf)  # the oracle's own parse of the same text
    out = _Inputs()
    out.oracle_catalog = ocat
    for teff, logg, nl, mh, w0, w1 in INSTANCES:
        atm = S.build_bundle(S.make_atm(teff, logg, nl, mh))
        out.append((atm, cat, w0, w1))
    return out


def _single(atm, cat, w0, w1):
    import torch

    from kurucz_py_joss_submission_amd import lines as L
    from kurucz_py_joss_submission_amd.device import SpectrumPlan

    plan = SpectrumPlan(atm, L.filter_by_range(cat, w0, w1), w0, w1, 300000.0)
    plan.run()
    torch.cuda.synchronize()
    return plan.flux.cpu().numpy(), plan.cont.cpu().numpy()


@pytest.mark.gpu
def test_batch_graph_equals_single_plans(gpu, batch_inputs):
    import torch

    from kurucz_py_joss_submission_amd.device import BatchPlan

    ref = [_single(*x) for x in batch_inputs]
    bp = BatchPlan(batch_inputs, autotune=True, graph=True)
    info = bp.info()
    assert info["pixels_per_lane"] in (2, 4, 8) and len(info["autotune_ms"]) == 4
    for replay in range(3):
        bp.run()
        torch.cuda.synchronize()
        got = bp.results()
        for i, (f, c) in enumerate(ref):
            assert got[i]["flux"].shape == f.shape
            assert np.array_equal(got[i]["flux"], f), (replay, i)
            assert np.array_equal(got[i]["continuum"], c), (replay, i)
        # eager work between replays
        _single(*batch_inputs[replay % len(batch_inputs)])
    bp.close()


@pytest.mark.gpu
def test_batch_eager_equals_graph(gpu, batch_inputs):
    import torch

    from kurucz_py_joss_submission_amd.device import BatchPlan

    a = BatchPlan(batch_inputs, autotune=False, graph=False)
    a.run()
    b = BatchPlan(batch_inputs, autotune=False, graph=True)
    b.run()
    torch.cuda.synchronize()
    for x, y in zip(a.results(), b.results()):
        assert np.array_equal(x["flux"], y["flux"]) and np.array_equal(x["continuum"], y["continuum"])
    a.close()
    b.close()


@pytest.mark.gpu
def test_batch_first_instance_against_oracle(gpu, oracle, batch_inputs):
    import torch

    from kurucz_py_joss_submission_amd import lines as L
    from kurucz_py_joss_submission_amd.device import BatchPlan

    atm, cat, w0, w1 = batch_inputs[2]
    bp = BatchPlan([batch_inputs[2]], autotune=False, graph=True)
    bp.run()
    torch.cuda.synchronize()
    r = bp.results()[0]
    import catalog as C

    ref = oracle.synthesize(atm, C.filter_by_range(batch_inputs.oracle_catalog, w0, w1), w0, w1, 300000.0)
    floor = 1e-3 * np.median(ref.continuum)
    assert np.max(np.abs(r["continuum"] - ref.continuum) / np.abs(ref.continuum)) < 1e-6
    assert np.max(np.abs(r["flux"] - ref.flux) / np.maximum(np.abs(ref.flux), floor)) < 1e-6
    bp.close()
