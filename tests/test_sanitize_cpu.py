"""The oracle and the CPU builds of the device headers under AddressSanitizer
and UndefinedBehaviorSanitizer.

tests/native/sanitize_main.cpp drives every oracle entry point (tables,
Voigt, FASTEX, PARCOE/INTEG/DERIV/MAP1, the JOSH column in both precisions,
continuum, frequency quantities, TRANSP, wing parameters, ASYNTH, the LTE
spectrum with and without line scattering, F_lambda, the hydrogen states,
profiles and walk, including lines off the grid and a refused n_upper) and the queue-record,
streaming-RT and RT-adjoint harnesses on small seeded inputs, sized as the
oracle's Python wrapper sizes them, including lines at and beyond the window's
ends and out-of-range element and ion indices.  The executable is built with
-fsanitize=address,undefined -fno-sanitize-recover=all, so any out-of-bounds
access, use after free, leak or undefined operation ends it with a report and
a non-zero status.
"""

from __future__ import annotations

import os
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
NATIVE = ROOT / "tests" / "native"
INC = ROOT / "kurucz_py_joss_submission_amd" / "csrc"
FLAGS = ["-O1", "-g", "-ffp-contract=off", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
         "-fno-omit-frame-pointer"]


def _have_sanitizers() -> bool:
    gcc = shutil.which("gcc")
    if gcc is None:
        return False
    lib = subprocess.run([gcc, "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    return os.path.isabs(lib) and Path(lib).exists()


@pytest.mark.skipif(not _have_sanitizers(), reason="gcc with libasan not available")
def test_oracle_and_harnesses_are_sanitizer_clean(tmp_path):
    objs = []
    for src in ("synthe_oracle.c", "hydrogen_oracle.c"):
        obj = tmp_path / (src[:-2] + ".o")
        subprocess.run(["gcc", "-std=c11", "-fopenmp", *FLAGS, "-c", str(ROOT / "oracle" / src), "-o", str(obj)],
                       check=True)
        objs.append(obj)
    for name in ("sanitize_main", "gather_queue_check", "rt_core_check", "rt_adj_check"):
        obj = tmp_path / f"{name}.o"
        subprocess.run(["g++", "-std=c++17", *FLAGS, f"-I{INC}", "-c", str(NATIVE / f"{name}.cpp"), "-o", str(obj)],
                       check=True)
        objs.append(obj)
    exe = tmp_path / "sanitize_main"
    subprocess.run(["g++", "-fsanitize=address,undefined", "-fopenmp", *map(str, objs), "-o", str(exe), "-lm"],
                   check=True)

    z = np.load(INC.parent / "data" / "kurucz_tables.npz")
    tables = tmp_path / "tables.bin"
    with open(tables, "wb") as f:
        np.array([z["tabvi"].size], np.int64).tofile(f)
        for k in ("tabvi", "tabh1", "xtau", "ch", "coefj"):
            np.ascontiguousarray(z[k], np.float64).tofile(f)
        for k in ("h_propbm", "h_c", "h_d", "h_pp", "h_beta_grid", "h_stalph", "h_stwtal", "h_istal", "h_lnghal",
                  "h_stcomp", "h_stcpwt", "h_lncomp", "h_cutoff_h2_plus", "h_cutoff_h2", "h_asum_lyman", "h_asum",
                  "h_y1wtm", "h_xknmtb", "conth", "ehyd_cm"):
            a = np.ascontiguousarray(z[k], np.float64)
            np.array([a.size], np.int64).tofile(f)
            a.tofile(f)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1", OMP_NUM_THREADS="2")
    r = subprocess.run([str(exe), str(tables)], capture_output=True, text=True, env=env, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "sanitize_main: done" in r.stdout
    assert "runtime error" not in r.stderr and "ERROR: AddressSanitizer" not in r.stderr, r.stderr[-4000:]
    assert r.stdout.count("oracle spectrum") == 3
    assert "hydrogen rc=0 rc_bad=-1" in r.stdout
