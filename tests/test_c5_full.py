"""Batched inverse problem at full size: 64 target spectra (atmospheres drawn
as the batched configuration, 5,000-point windows in 400-700 nm, 30,000 lines
each, all 19 non-hydrogen abundances offset by N(0, 0.2) dex; targets from the
FP64 plain path), recovered by Gauss-Newton from zero offsets over one
BatchPlan graph.

The misfit counts the pixels where neither target nor model takes the JOSH
label-401 branch (a saturated pixel crossing that switch jumps by about the
median flux).  Checked: the total misfit falls by five orders of magnitude and
every target whose starting misfit exceeds the numerical floor (1e-9) by two;
the constrained line elements come back (median error below 1e-3 dex); and the
FP64 plain path, run independently at the recovered offsets, reproduces the
targets to 1e-5 on the fitted pixels; the batched adjoint's gradient of the
misfit equals the per-spectrum differentiable model's.
"""

from __future__ import annotations

import numpy as np
import pytest


@pytest.mark.gpu
def test_c5_inverse_problem_full_size(gpu):
    from kurucz_py_joss_submission_amd import inverse as INV

    targets = INV.make_targets(n=64, n_lines=30_000, n_wl=5_000, seed=5)
    prob = INV.InverseProblem(targets)
    zero = np.zeros((prob.n, prob.nz))
    c0 = prob._costs(prob.residuals(prob.evaluate(zero)), prob.valid)
    sol = prob.solve(max_iter=20)
    cf = sol["cost"]
    rep = INV.report(targets, sol["delta"], sol["sensitivity"], marginal=sol["marginal"])
    live = c0 > 1e-9
    print(f"misfit {c0.sum():.4g} -> {cf.sum():.4g}; worst ratio over {int(live.sum())} live targets "
          f"{np.max(cf[live] / c0[live]):.3g}; {rep}")
    assert cf.sum() < 1e-5 * c0.sum()
    assert np.all(cf[live] < 1e-2 * c0[live])
    assert rep["line_pairs_constrained"] >= 40 and rep["median_err_constrained_line_dex"] < 1e-3
    # constrained by the marginal sensitivity: no element pair that trades off against another
    assert rep["max_err_constrained_line_dex"] < 0.05
    order = np.argsort(-cf / np.maximum(c0, 1e-300))
    check = INV.fp64_plain_check(targets, sol["delta"], indices=list(range(4)) + order[:4].tolist())
    print("FP64 plain path at the recovered offsets:", check)
    assert check < 1e-5
    # the batched adjoint at full size, at the recovered offsets: every spectrum's gradient of the
    # misfit against the per-spectrum differentiable model, on two spectra
    import torch

    from kurucz_py_joss_submission_amd.adjoint import SpectrumModel

    res = prob.residuals(prob.evaluate(sol["delta"]))
    masks = list(prob.valid)
    g = prob.gradient(res, masks)
    for i in (int(order[0]), int(np.argmax(sol["sensitivity"].sum(axis=1)))):
        t = targets[i]
        m = SpectrumModel(t.cols, t.catalog, t.wl_start, t.wl_end, precision="fp64")
        T, _, lgf, lgw = m.base_parameters()
        da = torch.tensor(sol["delta"][i], dtype=torch.float64, device="cuda", requires_grad=True)
        r = masks[i].to(torch.float64) * (m(T, da, lgf, lgw) - prob.f_target[i]) / prob.f_scale[i]
        (0.5 * (r * r).sum()).backward()
        ref = da.grad.cpu().numpy()
        print("batched adjoint vs model, spectrum", i, np.abs(g[i] - ref).max() / np.abs(ref).max())
        assert np.abs(g[i] - ref).max() <= 1e-6 * np.abs(ref).max()
    prob.plan.close()
