"""Batched Gauss-Newton abundance recovery on the GPU (small form of the
64-spectrum inverse problem): noise-free targets from the FP64 plain path at
random offsets N(0, 0.2) dex, recovered from zero offsets by the batched
forward pass.  The line elements (Fe, Ti, Cr, Ca, Ni) must come back to
1e-3 dex; the recovered spectra must match the targets under the FP64 plain
path; and the batched Jacobian must agree with the adjoint's gradient."""

from __future__ import annotations

import numpy as np
import pytest


@pytest.fixture(scope="module")
def small_problem(gpu):
    from kurucz_py_joss_submission_amd.inverse import InverseProblem, make_targets

    targets = make_targets(n=4, n_lines=3000, n_wl=800, seed=11)
    prob = InverseProblem(targets, autotune=False)
    sol = prob.solve(max_iter=10)
    return targets, prob, sol


@pytest.mark.gpu
def test_gauss_newton_recovers_line_elements(small_problem):
    """Every spectrum is fitted to 1e-6 of its starting misfit; every line
    element a spectrum constrains comes back to 2e-3 dex (the hot targets'
    neutral-metal lines are too weak to constrain theirs)."""
    from kurucz_py_joss_submission_amd.inverse import report

    targets, prob, sol = small_problem
    rep = report(targets, sol["delta"], sol["sensitivity"], marginal=sol["marginal"])
    print(sol["iterations"], sol["evaluations"], sol["cost"], rep)
    c0 = prob._costs(prob.residuals(prob.evaluate(np.zeros_like(sol["delta"]))), prob.valid)
    assert np.all(sol["cost"] < 1e-6 * c0)
    assert rep["line_pairs_constrained"] >= 5
    assert rep["max_err_constrained_line_dex"] < 2e-3


@pytest.mark.gpu
def test_recovered_spectra_under_fp64_plain_path(small_problem):
    from kurucz_py_joss_submission_amd.inverse import fp64_plain_check

    targets, prob, sol = small_problem
    assert fp64_plain_check(targets, sol["delta"], indices=[0, 1]) < 1e-4


@pytest.mark.gpu
def test_batched_jacobian_against_adjoint(small_problem):
    """d(sum w F)/d(offset) of spectrum 0: central differences of the batched
    forward (2e-3 dex) against the hand-written adjoint (SpectrumModel,
    FP64 profiles) for the line elements, to 1 %: the forward pass is
    piecewise smooth (KAPMIN gates, table steps, float32 JOSH arrays), which
    the difference sees and the adjoint, holding branches fixed, does not."""
    import torch

    from kurucz_py_joss_submission_amd import synthetic as S
    from kurucz_py_joss_submission_amd.adjoint import SpectrumModel
    from kurucz_py_joss_submission_amd.inverse import LINE_ELEMENTS

    targets, prob, sol = small_problem
    t = targets[0]
    delta = sol["delta"].copy()
    m = SpectrumModel(t.cols, t.catalog, t.wl_start, t.wl_end, precision="fp64")
    T, da, lgf, lgw = m.base_parameters()
    da = torch.tensor(delta[0], dtype=torch.float64, device="cuda", requires_grad=True)
    f = m(T, da, lgf, lgw)
    rng = np.random.default_rng(2)
    w = torch.tensor(rng.uniform(0.5, 1.5, m.W), dtype=torch.float64, device="cuda")
    (w * f).sum().backward()
    g_adj = da.grad.cpu().numpy()
    h = 1e-3
    for z in LINE_ELEMENTS:
        k = S.ABUNDANCE_Z.index(z)
        up, dn = delta.copy(), delta.copy()
        up[:, k] += h
        dn[:, k] -= h
        fu = prob.evaluate(up)[0]
        fd = prob.evaluate(dn)[0]
        g_fd = float((w * (fu - fd)).sum()) / (2 * h)
        print(z, g_adj[k], g_fd)
        assert abs(g_adj[k] - g_fd) <= 1e-2 * abs(g_fd) + 1e-9 * np.abs(g_adj).max()


@pytest.mark.gpu
def test_batched_adjoint_gradient_matches_spectrum_model(small_problem):
    """InverseProblem.gradient (the RT and line adjoints run on the batch's own
    forward arrays, chained through the closed-form bundle response) against
    the per-spectrum differentiable model (SpectrumModel, FP64 profiles, its
    torch layer model) for L = 1/2 sum (m r)^2, every abundance."""
    import torch

    from kurucz_py_joss_submission_amd.adjoint import SpectrumModel

    targets, prob, sol = small_problem
    rng = np.random.default_rng(4)
    delta = sol["delta"] + rng.normal(0.0, 0.02, sol["delta"].shape)
    delta[:, prob.h_index] = 0.0
    res = prob.residuals(prob.evaluate(delta))
    masks = list(prob.valid)
    g = prob.gradient(res, masks)
    for i, t in enumerate(targets):
        m = SpectrumModel(t.cols, t.catalog, t.wl_start, t.wl_end, precision="fp64")
        T, _, lgf, lgw = m.base_parameters()
        da = torch.tensor(delta[i], dtype=torch.float64, device="cuda", requires_grad=True)
        f = m(T, da, lgf, lgw)
        ft = torch.tensor(t.flux, dtype=torch.float64, device="cuda")
        r = masks[i].to(torch.float64) * (f - ft) / prob.f_scale[i]
        (0.5 * (r * r).sum()).backward()
        ref = da.grad.cpu().numpy()
        scale = np.abs(ref).max()
        print(i, np.abs(g[i] - ref).max() / scale)
        assert scale > 0
        assert np.abs(g[i] - ref).max() <= 1e-6 * scale


@pytest.mark.gpu
def test_adjoint_gauss_newton_recovers_line_elements(small_problem):
    """The same problem solved with the batched adjoint's gradient and BFGS
    updates of the Gauss-Newton Hessian (the difference Jacobian recomputed
    only after failed trials): the same bars as the difference solver, with
    fewer batched forward evaluations."""
    from kurucz_py_joss_submission_amd.inverse import InverseProblem, report

    targets, prob, sol = small_problem
    p2 = InverseProblem(targets, autotune=False)
    sol2 = p2.solve(max_iter=40, gradient="adjoint")
    rep = report(targets, sol2["delta"], sol2["sensitivity"], marginal=sol2["marginal"])
    print(sol2["iterations"], sol2["evaluations"], sol2["adjoint_evaluations"], sol2["jacobians"], sol2["cost"], rep,
          "differences:", sol["evaluations"])
    np.set_printoptions(linewidth=200, precision=3)
    print(sol2["cost_per_target"])
    c0 = p2._costs(p2.residuals(p2.evaluate(np.zeros_like(sol2["delta"]))), p2.valid)
    assert np.all(sol2["cost"] < 1e-6 * c0)
    assert rep["line_pairs_constrained"] >= 5
    assert rep["max_err_constrained_line_dex"] < 2e-3
    assert sol2["adjoint_evaluations"] >= 1
    assert sol2["evaluations"] < sol["evaluations"]
