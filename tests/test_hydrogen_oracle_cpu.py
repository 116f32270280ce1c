"""The oracle's HPROF4 hydrogen profile (oracle/hydrogen_oracle.c) against
the reference's own values and against what the mathematics fixes.

* tests/golden/hprof_reference.npz (tests/golden/make_hprof_fixture.py): the
  reference's hydrogen_line_profile on 22 transitions (Lyman alpha to Pfund
  and high series members) x 12 depth states x 10 offsets from the core to
  the far wing, SOFBETA on 407 (beta, p, n, m) including its branch
  boundaries, VCSE1F on 307 arguments including its branch boundaries;
* VCSE1F approximates the exponential integral E1 (Abramowitz & Stegun
  5.1.11, 5.1.53, 5.1.56): the series' first terms below x = 0.01, within
  2e-7 absolute up to x = 1 and the rational form (times FASTEX) above;
* HFNM approximates the hydrogen absorption oscillator strengths (Menzel &
  Pekeris): Lyman alpha 0.4162, H alpha 0.6407, H beta 0.1193, Paschen
  alpha 0.8421 within 1%;
* the fine-structure component weights of every (n, m) sum to one;
* SOFBETA's far branch is the Holtsmark asymptote 1.5 beta^-5/2 (1 + 18
  beta^-3/2) times its correction, which tends to one.
"""

from __future__ import annotations

from pathlib import Path
from types import SimpleNamespace

import numpy as np
import pytest

FIX = Path(__file__).resolve().parent / "golden" / "hprof_reference.npz"


def _states(oracle, inputs):
    T, xne, he1, h2, p0, p1, dopph = (inputs[:, i].copy() for i in range(7))
    D = T.size
    atm = SimpleNamespace(layers=D, temperature=T, electron_density=xne, xnf_he1=he1, xnf_h2=h2,
                          xnfph=np.stack([p0, p1], axis=1), dopph=dopph, turbulent_velocity=np.zeros(D))
    return oracle.hydrogen_states(atm)


def test_profile_matches_reference_values(oracle):
    z = np.load(FIX)
    st = _states(oracle, z["state_inputs"])
    worst = 0.0
    for n, m, si, dl, ref in z["profile"]:
        got = oracle.hydrogen_profile(int(n), int(m), st[int(si)], float(dl))
        if ref == 0.0:
            assert got == 0.0
            continue
        worst = max(worst, abs(got - ref) / abs(ref))
    assert worst < 1e-13, worst


def test_sofbeta_and_vcse1f_match_reference(oracle):
    z = np.load(FIX)
    for beta, pp, n, m, ref in z["sofbeta"]:
        got = oracle.sofbeta(float(beta), float(pp), int(n), int(m))
        assert got == pytest.approx(ref, rel=1e-14, abs=0.0), (beta, pp, n, m)
    for x, ref in z["vcse1f"]:
        assert oracle.vcse1f(float(x)) == pytest.approx(ref, rel=1e-14, abs=1e-300), x


def test_vcse1f_is_e1(oracle):
    from scipy.special import exp1

    for x in np.concatenate([np.geomspace(1e-4, 1.0, 40), np.geomspace(1.0001, 30.0, 40)]):
        v, e = oracle.vcse1f(float(x)), float(exp1(x))
        if x <= 0.01:  # -ln x - 0.577215 + x: the truncated Euler constant and the dropped x^2/4
            assert abs(v - e) < 7e-7 + x * x / 4, x
        elif x <= 1.0:
            assert abs(v - e) < 2.5e-7, x  # A&S 5.1.53: 2e-7, plus the truncated Euler constant
        else:
            # the rational form times FASTEX, whose table carries exp(-x) to ~5e-4
            assert abs(v - e) / e < 6e-4, x
    assert oracle.vcse1f(31.0) == 0.0 and oracle.vcse1f(0.0) == 0.0


def test_hf_nm_oscillator_strengths(oracle):
    for (n, m), f in {(1, 2): 0.4162, (2, 3): 0.6407, (2, 4): 0.1193, (3, 4): 0.8421, (1, 3): 0.0791}.items():
        assert oracle.hf_nm(n, m) == pytest.approx(f, rel=0.01), (n, m)
    assert oracle.hf_nm(3, 3) == 0.0 and oracle.hf_nm(4, 2) == 0.0


def test_fine_structure_weights_sum_to_one(oracle):
    for n in range(1, 6):
        for m in range(n + 1, 13):
            off, wt = oracle.fine_structure(n, m)
            assert off.size == wt.size >= 1
            if n > 4 or m > 10:
                assert off.tolist() == [0.0] and wt.tolist() == [1.0]
            elif m - n != 1 or n < 4:
                assert wt.sum() == pytest.approx(1.0, abs=1e-3), (n, m)
    # Brackett alpha: the reference's 0-based slice from the 1-based start
    # (hydrogen.py:1530-1532) runs off the table's end and keeps 13 of 14
    off, wt = oracle.fine_structure(4, 5)
    assert off.size == 13


def test_sofbeta_far_branch_is_holtsmark(oracle):
    for beta in (600.0, 2e3, 1e4, 1e5):
        h = 1.5 * beta ** -2.5 * (1.0 + 18.0 * beta ** -1.5)
        assert oracle.sofbeta(beta, 0.3, 2, 3) == pytest.approx(h, rel=1e-14)
    # the near branch's correction tends to one with the Holtsmark form at 25.12
    lo, hi = oracle.sofbeta(25.12, 0.3, 2, 3), oracle.sofbeta(25.1201, 0.3, 2, 3)
    assert hi == pytest.approx(lo, rel=0.05)
