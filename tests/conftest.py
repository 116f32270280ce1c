"""Shared fixtures: the oracle (CPU checker), seeded workloads, GPU gating.

Workloads (atmosphere bundle + parsed line list) are cached for the session
in a temporary directory, or in SYNTHE_TEST_CACHE when that names a
directory the caller keeps between sessions.
"""

from __future__ import annotations

import os
import shutil
import sys
import tempfile
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))

_OWN_CACHE = not os.environ.get("SYNTHE_TEST_CACHE")
CACHE = Path(os.environ["SYNTHE_TEST_CACHE"]) if not _OWN_CACHE else Path(tempfile.mkdtemp(prefix="synthe_wk_"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and libsynthe_hip.so")


def pytest_unconfigure(config):
    if _OWN_CACHE:
        shutil.rmtree(CACHE, ignore_errors=True)


@pytest.fixture(scope="session")
def oracle():
    import oracle as O

    O.build()
    return O


def make_small(name="small", teff=5750.0, logg=4.44, n_lines=2000, wl0=500.0, wl1=501.0, seed=7,
               n_layers=72, mh=0.0, line_wl=None):
    from kurucz_py_joss_submission_amd import synthetic as s

    return s.make_workload(name, teff, logg, n_lines, wl0, wl1, seed=seed, n_layers=n_layers,
                           cache_dir=CACHE, mh=mh, line_wl=line_wl)


def oracle_catalog(wk, window=None):
    """The oracle's own parse of a workload's GFALL text (catalog.read_gfall),
    DELLIM-filtered to window (default: the workload's window)."""
    import catalog as C
    from kurucz_py_joss_submission_amd import synthetic as s

    w0, w1 = window if window is not None else (wk.wl_start, wk.wl_end)
    with tempfile.TemporaryDirectory() as td:
        path = s.write_workload_gfall(wk, Path(td) / "lines.gfall", window=(w0, w1))
        return C.filter_by_range(C.read_gfall(path), w0, w1)


@pytest.fixture(scope="session")
def small_workload():
    return make_small()


@pytest.fixture(scope="session")
def c1_workload():
    """Config 1: Teff 5750, log g 4.44, 72 layers, 20,000 lines over 500-510 nm."""
    return make_small("c1", 5750.0, 4.44, 20000, 500.0, 510.0, seed=1)


@pytest.fixture(scope="session")
def gpu():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from kurucz_py_joss_submission_amd import _native

    return _native.context()
