"""The RT adjoint (csrc/rt_adjoint.h) compiled for the CPU.

1. Its FP64 total-flux model agrees with the oracle's solve_josh_flux
   (physics/josh_solver.py:367-767, float32 sweep to 1e-5) to that tolerance.
2. Its vector-Jacobian product equals central finite differences of that
   model in every input (abtot, alpha, snubar per depth), for the XTAU
   branch and the label-401 branch.
"""

from __future__ import annotations

import ctypes as ct
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
SRC = ROOT / "tests" / "native" / "rt_adj_check.cpp"
INC = ROOT / "kurucz_py_joss_submission_amd" / "csrc"
TABLES = ROOT / "kurucz_py_joss_submission_amd" / "data" / "kurucz_tables.npz"


@pytest.fixture(scope="module")
def adj(tmp_path_factory):
    out = tmp_path_factory.mktemp("adj") / "libadj.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
                    f"-I{INC}", str(SRC), "-o", str(out)], check=True)
    lib = ct.CDLL(str(out))
    vp = ct.c_void_p
    lib.adj_flux64.restype = ct.c_double
    lib.adj_flux64.argtypes = [ct.c_int] + [vp] * 7 + [ct.c_double, ct.POINTER(ct.c_int)]
    lib.adj_vjp.restype = ct.c_double
    lib.adj_vjp.argtypes = [ct.c_int] + [vp] * 7 + [ct.c_double] + [vp] * 3
    z = np.load(TABLES)
    tabs = (np.ascontiguousarray(z["xtau"], np.float64), np.ascontiguousarray(z["ch"], np.float64),
            np.ascontiguousarray(z["coefj"], np.float32))
    return lib, tabs


def _p(a):
    return a.ctypes.data


def _column(rng, n, thick=False):
    """A smooth column like a stellar atmosphere's: opacity rising with column
    mass (power law, mild noise), scattering albedo 1e-3..0.5, a source
    function rising with depth."""
    rho = np.logspace(rng.uniform(-5, -4), rng.uniform(1, 1.7), n)
    k0 = 10 ** rng.uniform(-1.5, 0.5) * (1e8 if thick else 1.0)
    acont = k0 * (rho / rho[0]) ** rng.uniform(0.3, 0.8) * (1 + 0.05 * rng.standard_normal(n))
    albedo = 10 ** rng.uniform(-3, -0.3, n)
    sigmac = acont * albedo / (1 - albedo)
    scont = 1e-5 * (1 + (rho / rho[-1]) ** 0.25 * rng.uniform(2, 6)) * (1 + 0.02 * rng.standard_normal(n))
    abtot = acont + sigmac
    return rho, abtot, sigmac / abtot, scont, acont, sigmac


def _flux(lib, tabs, rho, ab, al, sb, tol=1e-13):
    t = ct.c_int(0)
    f = lib.adj_flux64(len(rho), _p(rho), _p(ab), _p(al), _p(sb), *map(_p, tabs), tol, ct.byref(t))
    return f, t.value


@pytest.mark.parametrize("seed", range(6))
def test_fp64_model_matches_oracle(adj, oracle, seed):
    lib, tabs = adj
    rng = np.random.default_rng(seed)
    rho, ab, al, sb, acont, sigmac = _column(rng, int(rng.integers(30, 80)))
    f, t401 = _flux(lib, tabs, rho, ab, al, sb)
    z = np.zeros_like(rho)
    ref = oracle.solve_josh_flux(acont, sb, z, sb, sigmac, z, rho)[0]
    assert t401 == 0
    assert abs(f / ref - 1) < 3e-5


@pytest.mark.parametrize("seed,thick", [(s, False) for s in range(5)] + [(s, True) for s in range(3)])
def test_vjp_against_central_differences(adj, seed, thick):
    lib, tabs = adj
    rng = np.random.default_rng(100 + seed)
    n = int(rng.integers(20, 60))
    rho, ab, al, sb, _, _ = _column(rng, n, thick)
    g = [np.zeros(n) for _ in range(3)]
    v = 1.7
    f = lib.adj_vjp(n, _p(rho), _p(ab), _p(al), _p(sb), *map(_p, tabs), v, *map(_p, g))
    f0, t401 = _flux(lib, tabs, rho, ab, al, sb)
    assert t401 == int(thick) and f == f0
    fds = []
    for q, arr in enumerate((ab, al, sb)):
        fd = np.zeros(n)
        for j in range(n):
            h = 1e-6 * arr[j]
            up, dn = arr.copy(), arr.copy()
            up[j] += h
            dn[j] -= h
            args = [ab, al, sb]
            args[q] = up
            fu, tu = _flux(lib, tabs, rho, *args)
            args[q] = dn
            fdn, td = _flux(lib, tabs, rho, *args)
            assert tu == td == t401
            fd[j] = v * (fu - fdn) / (2 * h)
        fds.append(fd)
    # every component against the column's largest gradient (inputs scaled by their values)
    scale = max(np.max(np.abs(fd * arr)) for fd, arr in zip(fds, (ab, al, sb)))
    for q, (fd, gq, arr) in enumerate(zip(fds, g, (ab, al, sb))):
        err = np.max(np.abs((gq - fd) * arr)) / scale
        assert err < 2e-5, (q, err)


@pytest.mark.parametrize("thick", [False, True])
def test_reverse_vjp_matches_forward_mode_by_inputs(adj, thick):
    """The O(n) vjp (colour-seeded Jacobians of PARCOE/MAP1, implicit label-401
    adjoint) against forward-mode dual numbers seeded one input at a time
    (O(n^2), the definition).  The XTAU branch contracts the same partial
    derivatives (equal to rounding); the label-401 branch compares the implicit
    derivative with the derivative of the iteration run to 1e-12."""
    lib, tabs = adj
    vp = ct.c_void_p
    lib.adj_vjp_by_inputs.restype = ct.c_double
    lib.adj_vjp_by_inputs.argtypes = [ct.c_int] + [vp] * 7 + [ct.c_double] + [vp] * 3
    tol = 1e-8 if thick else 1e-13
    for seed in range(8):
        rng = np.random.default_rng(500 + seed)
        n = int(rng.integers(5, 96))
        rho, ab, al, sb, _, _ = _column(rng, n, thick)
        g = [np.zeros(n) for _ in range(3)]
        h = [np.zeros(n) for _ in range(3)]
        f1 = lib.adj_vjp(n, _p(rho), _p(ab), _p(al), _p(sb), *map(_p, tabs), 0.7, *map(_p, g))
        f2 = lib.adj_vjp_by_inputs(n, _p(rho), _p(ab), _p(al), _p(sb), *map(_p, tabs), 0.7, *map(_p, h))
        assert f1 == f2
        scale = max(np.max(np.abs(b * x)) for b, x in zip(h, (ab, al, sb)))
        for q, (a, b, x) in enumerate(zip(g, h, (ab, al, sb))):
            err = np.max(np.abs((a - b) * x)) / scale
            assert err < tol, (seed, n, q, err)


@pytest.mark.parametrize("seed", range(3))
def test_vjp_strong_scattering(adj, seed):
    """Albedo 0.995-0.9995 at every depth (electron scattering in hot,
    low-gravity atmospheres): Gauss-Seidel contracts by ~0.999 per sweep, and
    the Krylov finish must bring the fixed point and its adjoint to
    convergence, so that the vjp equals differences of the converged model."""
    lib, tabs = adj
    rng = np.random.default_rng(700 + seed)
    n = 60
    rho = np.logspace(-4, 1.5, n)
    acont = 0.3 * (rho / rho[0]) ** 0.5
    albedo = 1 - 10 ** rng.uniform(-3.3, -2.3, n)
    sig = acont * albedo / (1 - albedo)
    ab, al = acont + sig, sig / (acont + sig)
    sb = 1e-5 * (1 + (rho / rho[-1]) ** 0.25 * 4)
    f_loose, _ = _flux(lib, tabs, rho, ab, al, sb, tol=1e-10)
    f_tight, t401 = _flux(lib, tabs, rho, ab, al, sb, tol=1e-15)
    assert t401 == 0 and abs(f_loose / f_tight - 1) < 1e-12
    g = [np.zeros(n) for _ in range(3)]
    lib.adj_vjp(n, _p(rho), _p(ab), _p(al), _p(sb), *map(_p, tabs), 1.0, *map(_p, g))
    fds = []
    for q, arr in enumerate((ab, al, sb)):
        fd = np.zeros(n)
        for j in range(n):
            h = 1e-7 * arr[j]
            args_u, args_d = [ab.copy(), al.copy(), sb.copy()], [ab.copy(), al.copy(), sb.copy()]
            args_u[q][j] += h
            args_d[q][j] -= h
            fd[j] = (_flux(lib, tabs, rho, *args_u)[0] - _flux(lib, tabs, rho, *args_d)[0]) / (2 * h)
        fds.append(fd)
    scale = max(np.max(np.abs(fd * arr)) for fd, arr in zip(fds, (ab, al, sb)))
    for q, (fd, gq, arr) in enumerate(zip(fds, g, (ab, al, sb))):
        assert np.max(np.abs((gq - fd) * arr)) / scale < 1e-6, q
