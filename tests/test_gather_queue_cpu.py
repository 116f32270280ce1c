"""The wing gather's queue record (csrc/gather_queue.h), compiled for the CPU:
every field survives a pack/unpack at its bounds -- centres a MAXSTEP before
the tile or a MAXSTEP past its last pixel, half-widths up to the largest the
wing set-up lets through (MAXPROF = 10^6 pixels) -- and the host rejects
grids whose Voigt zones would pass MAXPROF."""

from __future__ import annotations

import ctypes as ct
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]


@pytest.fixture(scope="module")
def gq(tmp_path_factory):
    out = tmp_path_factory.mktemp("gq") / "libgq.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-Werror",
                    f"-I{ROOT / 'kurucz_py_joss_submission_amd' / 'csrc'}",
                    str(ROOT / "tests" / "native" / "gather_queue_check.cpp"), "-o", str(out)], check=True)
    lib = ct.CDLL(str(out))
    lib.gq_roundtrip.argtypes = [ct.c_int, ct.c_int, ct.c_int, ct.c_void_p]
    lib.gq_max_zone.restype = ct.c_long
    lib.gq_max_tile.restype = ct.c_long
    return lib


def _rt(lib, dc, r, m):
    out = np.zeros(3, dtype=np.int64)
    lib.gq_roundtrip(dc, r, m, out.ctypes.data)
    return out.tolist()


def test_fields_round_trip_at_their_bounds(gq):
    zmax, tile = gq.gq_max_zone(), gq.gq_max_tile()
    assert zmax == 1_000_000 and tile == 4096
    for dc in (-zmax, -zmax + 1, -1, 0, 1, 511, zmax + tile - 1, 2**20 - 1, -(2**20)):
        for r in (0, 1, 18, zmax - 1, zmax, 2**20 - 1):
            for m in (0, 1, 284, zmax, 2**20 - 1):
                assert _rt(gq, dc, r, m) == [dc, r, m]


def test_a_centre_past_the_field_would_not_survive(gq):
    """Why MAXSTEP is clamped to MAXPROF and tiles are bounded: 2^20 does not fit."""
    assert _rt(gq, 2**20, 5, 7) != [2**20, 5, 7]
    assert _rt(gq, 3, 2**20, 7) != [3, 2**20, 7]


def test_grids_past_maxprof_are_rejected_on_the_host():
    from kurucz_py_joss_submission_amd.device import check_wing_reach

    dop = np.full((3, 6, 30), 2e-6)  # 0.6 km/s
    g = 500.0 * (1 + 1 / 3e5) ** np.arange(10)
    check_wing_reach(g, dop)  # R = 3e5: N10DOP = 6
    fine = 500.0 * (1 + 1 / 6e10) ** np.arange(10)
    with pytest.raises(ValueError, match="MAXPROF"):
        check_wing_reach(fine, dop)  # N10DOP = 1.2e6
