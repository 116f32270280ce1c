"""The GPU RT core's streaming INTEG + MAP1, compiled for the CPU, against the oracle.

The device kernels replace the per-depth arrays of INTEG/PARCOE/MAP1
(physics/josh_solver.py:95-353) with a one-pass stream; this checks the
stream bit for bit against the oracle's array form on random columns.
"""

from __future__ import annotations

import ctypes as ct
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
SRC = ROOT / "tests" / "native" / "rt_core_check.cpp"
INC = ROOT / "kurucz_py_joss_submission_amd" / "csrc"


@pytest.fixture(scope="module")
def rtc(tmp_path_factory):
    out = tmp_path_factory.mktemp("rtc") / "librtc.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
                    f"-I{INC}", str(SRC), "-o", str(out)], check=True)
    lib = ct.CDLL(str(out))
    lib.rtc_integ_map1.restype = ct.c_double
    return lib


def _p(a):
    return a.ctypes.data_as(ct.POINTER(ct.c_double))


@pytest.mark.parametrize("seed", range(40))
def test_stream_matches_arrays(rtc, oracle, seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(4, 100))
    rho = np.sort(10 ** rng.uniform(-6, 2, n))
    if seed % 7 == 0:
        rho = np.unique(np.round(rho, 3)) + 1e-3  # near-ties in x
        n = rho.size
        if n < 4:
            return
    scale = 10 ** rng.uniform(-3, 5)
    abtot = scale * 10 ** rng.uniform(-2, 2, n)
    if seed % 5 == 0:
        abtot = np.full(n, scale)  # constant: PARCOE c == 0 everywhere
    snubar = 10 ** rng.uniform(2, 6, n)
    alpha = rng.uniform(0, 1, n) ** 3
    xs, xa = np.zeros(51), np.zeros(51)
    tau0 = rtc.rtc_integ_map1(n, _p(oracle.XTAU), _p(abtot), _p(rho), _p(snubar), _p(alpha),
                              _p(xs), _p(xa))
    assert tau0 >= 0.0
    taunu = oracle.integ(rho, abtot, abtot[0] * rho[0])
    assert tau0 == taunu[0]
    ref_s, _ = oracle.map1_josh(taunu, snubar, oracle.XTAU)
    ref_a, _ = oracle.map1_josh(taunu, alpha, oracle.XTAU)
    assert np.array_equal(xs, ref_s), np.max(np.abs(xs - ref_s) / np.abs(ref_s))
    assert np.array_equal(xa, ref_a)
