"""ctypes front end of the C oracle (TEST INFRASTRUCTURE ONLY).

The oracle is the FP64 CPU restatement of the reference hot path
(synthe_oracle.c).  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this module, as the checker or the timed CPU
baseline; the product path never does.

``synthesize`` runs the whole path on the CPU with the reference's stage
order (engine/opacity.py:2143-3861): grid, centre indices, continuum,
frequency quantities, TRANSP, ASYNTH, LTE source, RT, unit conversion.
"""

from __future__ import annotations

import ctypes as ct
import math
import os
import subprocess
from dataclasses import dataclass, field
from pathlib import Path
from typing import Dict, Optional

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB_PATH = _HERE / "liboracle.so"
_lib = None

_D = ct.POINTER(ct.c_double)
_F = ct.POINTER(ct.c_float)
_I64 = ct.POINTER(ct.c_int64)
_I32 = ct.POINTER(ct.c_int32)
_U8 = ct.POINTER(ct.c_uint8)
_INT = ct.POINTER(ct.c_int)


class _Tables(ct.Structure):
    _fields_ = [("h0", _D), ("h1", _D), ("h2", _D), ("nvoigt", ct.c_int),
                ("extab", _D), ("extabf", _D), ("nfastex", ct.c_int)]


def build(force: bool = False) -> Path:
    """Compile liboracle.so with the committed Makefile."""
    if force or not _LIB_PATH.exists() or _LIB_PATH.stat().st_mtime < max(f.stat().st_mtime for f in _HERE.glob("*.c")):
        subprocess.run(["make", "-C", str(_HERE), "-B" if force else "liboracle.so"], check=True,
                       stdout=subprocess.DEVNULL)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            build()
        L = ct.CDLL(str(_LIB_PATH))
        L.orc_voigt.restype = ct.c_double
        L.orc_fastex.restype = ct.c_double
        L.orc_wavelength_grid.restype = ct.c_long
        L.orc_solve_josh_flux.restype = ct.c_double
        L.orc_lte_frequency.restype = ct.c_double
        L.orc_map1_josh.restype = ct.c_int
        L.orc_josh_iterate.restype = ct.c_int
        L.orc_get_threads.restype = ct.c_int
        for name in ("orc_vcse1f", "orc_sofbeta", "orc_hf_nm", "orc_hydrogen_profile"):
            getattr(L, name).restype = ct.c_double
        L.orc_fine_structure.restype = ct.c_int
        L.orc_hydrogen_lines.restype = ct.c_int
        _lib = L
    return _lib


def p(a: np.ndarray, typ=_D):
    return a.ctypes.data_as(typ)


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def set_threads(n: int) -> None:
    lib().orc_set_threads(ct.c_int(int(n)))


def get_threads() -> int:
    return int(lib().orc_get_threads())


# ------------------------------------------------------------------ tables
_KT = np.load(Path(__file__).resolve().parents[1] / "kurucz_py_joss_submission_amd" / "data"
              / "kurucz_tables.npz")
XTAU = _KT["xtau"].copy()
CH = _KT["ch"].copy()
COEFJ = _KT["coefj"].copy()
TABVI = _KT["tabvi"].copy()
TABH1 = _KT["tabh1"].copy()


def voigt_tables(n: int = 2001):
    h0, h1, h2 = np.zeros(n), np.zeros(n), np.zeros(n)
    lib().orc_build_voigt_tables(p(TABVI), p(TABH1), ct.c_int(TABVI.size), p(h0), p(h1), p(h2),
                                 ct.c_int(n))
    return h0, h1, h2


def fastex_tables(n: int = 1001):
    e, f = np.zeros(n), np.zeros(n)
    lib().orc_build_fastex_tables(p(e), p(f), ct.c_int(n))
    return e, f


class Tables:
    """Owns the Voigt and FASTEX tables and the ctypes struct pointing at them."""

    def __init__(self):
        self.h0, self.h1, self.h2 = voigt_tables()
        self.extab, self.extabf = fastex_tables()
        self.struct = _Tables(p(self.h0), p(self.h1), p(self.h2), 2001, p(self.extab),
                              p(self.extabf), 1001)
        lib().orc_set_josh_tables(p(XTAU), p(CH), p(COEFJ))
        self.h = {k: np.ascontiguousarray(_KT[k], dtype=np.float64) for k in _KT.files
                  if k.startswith("h_") or k in ("conth", "ehyd_cm")}
        h = self.h
        lib().orc_set_hydrogen_tables(
            p(h["h_propbm"]), p(h["h_c"]), p(h["h_d"]), p(h["h_pp"]), p(h["h_beta_grid"]), p(h["h_stalph"]),
            p(h["h_stwtal"]), p(h["h_istal"]), p(h["h_lnghal"]), p(h["h_stcomp"]), p(h["h_stcpwt"]),
            p(h["h_lncomp"]), p(h["h_cutoff_h2_plus"]), ct.c_int(h["h_cutoff_h2_plus"].size),
            p(h["h_cutoff_h2"]), ct.c_int(h["h_cutoff_h2"].size), p(h["h_asum_lyman"]),
            ct.c_int(h["h_asum_lyman"].size), p(h["h_asum"]), ct.c_int(h["h_asum"].size), p(h["h_y1wtm"]),
            p(h["h_xknmtb"]), p(h["conth"]), ct.c_int(h["conth"].size), p(h["ehyd_cm"]), p(self.extab),
            p(self.extabf), ct.c_int(1001))


_TABLES: Optional[Tables] = None


def tables() -> Tables:
    global _TABLES
    if _TABLES is None:
        _TABLES = Tables()
    return _TABLES


def voigt(v: float, a: float) -> float:
    t = tables()
    return lib().orc_voigt(ct.c_double(v), ct.c_double(a), p(t.h0), p(t.h1), p(t.h2), ct.c_int(2001))


def fastex(x: float) -> float:
    t = tables()
    return lib().orc_fastex(ct.c_double(x), p(t.extab), p(t.extabf), ct.c_int(1001))


def map1_table(xold, fold, xnew):
    xold, fold, xnew = _c(xold, np.float64), _c(fold, np.float64), _c(xnew, np.float64)
    out = np.zeros_like(xnew)
    lib().orc_map1_table(p(xold), p(fold), ct.c_int(xold.size), p(xnew), p(out), ct.c_int(xnew.size))
    return out


def map1_josh(xold, fold, xnew):
    xold, fold, xnew = _c(xold, np.float64), _c(fold, np.float64), _c(xnew, np.float64)
    out = np.zeros_like(xnew)
    maxj = lib().orc_map1_josh(p(xold), p(fold), ct.c_int(xold.size), p(xnew), p(out), ct.c_int(xnew.size))
    return out, int(maxj)


def parcoe(f, x):
    f, x = _c(f, np.float64), _c(x, np.float64)
    a, b, c = np.zeros_like(f), np.zeros_like(f), np.zeros_like(f)
    lib().orc_parcoe(p(f), p(x), ct.c_long(f.size), p(a), p(b), p(c))
    return a, b, c


def integ(x, f, start):
    x, f = _c(x, np.float64), _c(f, np.float64)
    out = np.zeros_like(f)
    lib().orc_integ(p(x), p(f), ct.c_long(f.size), ct.c_double(start), p(out))
    return out


def deriv(x, f):
    x, f = _c(x, np.float64), _c(f, np.float64)
    out = np.zeros_like(f)
    lib().orc_deriv(p(x), p(f), ct.c_long(f.size), p(out))
    return out


def josh_iterate(xs, xalpha, xsbar_mod, max_iter=51):
    tables()
    xs = _c(xs, np.float32).copy()
    xa, xb = _c(xalpha, np.float32), _c(xsbar_mod, np.float32)
    it = lib().orc_josh_iterate(p(xs, _F), p(xa, _F), p(xb, _F), ct.c_int(max_iter))
    return xs, int(it)


def set_josh_fp64(on: bool, tol: float = 1e-13) -> None:
    """Test-only smooth mode: FP64 scattering sweep and label-401 iteration to tol."""
    lib().orc_set_josh_fp64(ct.c_int(int(on)), ct.c_double(tol))


def solve_josh_flux(acont, scont, aline, sline, sigmac, sigmal, rhox):
    tables()
    arrs = [_c(a, np.float64) for a in (acont, scont, aline, sline, sigmac, sigmal, rhox)]
    info = np.zeros(2, dtype=np.intc)
    f = lib().orc_solve_josh_flux(ct.c_long(arrs[0].size), *[p(a) for a in arrs], p(info, _INT))
    return float(f), int(info[0]), bool(info[1])


def wavelength_grid(start: float, end: float, resolution: float) -> np.ndarray:
    n = lib().orc_wavelength_grid(ct.c_double(start), ct.c_double(end), ct.c_double(resolution),
                                  None, ct.c_long(0))
    out = np.zeros(n)
    lib().orc_wavelength_grid(ct.c_double(start), ct.c_double(end), ct.c_double(resolution),
                              p(out), ct.c_long(n))
    return out


def nearest_grid_indices(grid, values):
    grid, values = _c(grid, np.float64), _c(values, np.float64)
    out = np.zeros(values.size, dtype=np.int64)
    lib().orc_nearest_grid_indices(p(grid), ct.c_long(grid.size), p(values), ct.c_long(values.size),
                                   p(out, _I64))
    return out


def wing_center_indices(grid, values, origin=None):
    grid, values = _c(grid, np.float64), _c(values, np.float64)
    out = np.zeros(values.size, dtype=np.int64)
    lib().orc_wing_center_indices(p(grid), ct.c_long(grid.size), p(values), ct.c_long(values.size),
                                  ct.c_double(origin if origin else 0.0), p(out, _I64))
    return out


def continuum(atm, grid):
    """build_depth_continuum on the CPU -> (abs[D,W], scat[D,W])."""
    D, W = atm.layers, grid.size
    wledge = _c(atm.continuum_wledge, np.float64)
    E = wledge.size
    half = atm.continuum_half_edge
    delta = atm.continuum_delta_edge
    half = _c(half, np.float64) if half is not None and half.size == E - 1 else None
    delta = _c(delta, np.float64) if delta is not None and delta.size == E - 1 else None
    ca, cs = _c(atm.continuum_abs_coeff, np.float64), _c(atm.continuum_scat_coeff, np.float64)
    out_a, out_s = np.zeros((D, W)), np.zeros((D, W))
    lib().orc_continuum(ct.c_long(D), ct.c_long(W), p(_c(grid, np.float64)), ct.c_long(E), p(wledge),
                        p(half) if half is not None else None, p(delta) if delta is not None else None,
                        p(ca), p(cs), ct.c_int(int(atm.continuum_coeff_log10)), p(out_a), p(out_s))
    return out_a, out_s


def flux_to_lambda(grid, flux_hz):
    """F_nu per Hz -> F_lambda per nm (engine/opacity.py:3852-3861)."""
    grid = _c(grid, np.float64)
    out = _c(flux_hz, np.float64).copy()
    lib().orc_flux_to_lambda(ct.c_long(grid.size), p(grid), p(out))
    return out


def freq_quantities(grid, temp):
    D, W = temp.size, grid.size
    e, s, b, sl = (np.zeros((D, W)) for _ in range(4))
    lib().orc_freq_quantities(ct.c_long(D), ct.c_long(W), p(_c(grid, np.float64)),
                              p(_c(temp, np.float64)), p(e), p(s), p(b), p(sl))
    return e, s, b, sl


def hckt(atm) -> np.ndarray:
    """HCKT per depth: the bundle's, else 1.4388/max(T, 1) (physics/populations.py:162-171)."""
    t = np.maximum(np.asarray(atm.temperature, np.float64), 1.0)
    out = 1.4388 / t
    h = getattr(atm, "hckt", None)
    if h is not None:
        h = np.asarray(h, np.float64)
        out[: min(h.size, t.size)] = h[: min(h.size, t.size)]
    return out


def txnxn(atm) -> np.ndarray:
    """TXNXN per depth: the bundle's, else (xnf_h + .42 xnf_he1 + .85 xnf_h2)(T/1e4)^.3
    with T = max(T, 1) (physics/populations.py:190-199)."""
    tx = getattr(atm, "txnxn", None)
    if tx is not None:
        return np.asarray(tx, np.float64)
    out = []
    for i, t in enumerate(np.asarray(atm.temperature, np.float64).tolist()):
        col = [getattr(atm, n, None) for n in ("xnf_h", "xnf_he1", "xnf_h2")]
        h, he, h2 = (float(c[i]) if c is not None else 0.0 for c in col)
        out.append((h + 0.42 * he + 0.85 * h2) * (max(t, 1.0) / 10_000.0) ** 0.3)
    return np.array(out, np.float64)


def _rho(atm):
    v = getattr(atm, "mass_density", None)  # ones when absent (physics/line_opacity.py:794)
    return np.ones(atm.temperature.size) if v is None else _c(v, np.float64)


def _xne(atm):
    v = getattr(atm, "electron_density", None)  # zeros when absent (physics/line_opacity.py:798)
    return np.zeros(atm.temperature.size) if v is None else _c(v, np.float64)


def transp(soa, atm, grid, kapmin, cutoff=1e-3, microturb_kms=0.0):
    t = tables()
    L, D, W = len(soa), atm.layers, grid.size
    pop = _c(atm.population_per_ion, np.float64)
    dop = _c(atm.doppler_per_ion, np.float64)
    out = np.zeros((L, D))
    valid = np.zeros((L, D), dtype=np.uint8)
    lib().orc_transp(
        ct.c_long(L), ct.c_long(D), ct.c_long(W), p(soa.process, _U8), p(soa.elem, _I32),
        p(soa.ion, _I32), p(soa.ltype, _I32), p(soa.wl), p(soa.gf), p(soa.cgf),
        p(soa.gamma_rad), p(soa.gamma_stark), p(soa.gamma_vdw), p(soa.elo), p(soa.center_t, _I64),
        p(hckt(atm)), p(pop), p(dop), ct.c_long(pop.shape[1]),
        ct.c_long(pop.shape[2]), p(_rho(atm)),
        p(_xne(atm)), p(txnxn(atm)),
        p(_c(kapmin, np.float64)), ct.c_double(cutoff), ct.c_double(microturb_kms),
        ct.byref(t.struct), p(out), p(valid, _U8))
    return out, valid.astype(bool)


def asynth(soa, atm, grid, transp_arr, valid, kapmin, cutoff=1e-3):
    t = tables()
    L, D, W = len(soa), atm.layers, grid.size
    pop = _c(atm.population_per_ion, np.float64)
    dop = _c(atm.doppler_per_ion, np.float64)
    out = np.zeros((D, W))
    stats = np.zeros(5)
    lib().orc_asynth(
        ct.c_long(L), ct.c_long(D), ct.c_long(W), p(_c(grid, np.float64)), p(soa.ltype, _I32),
        p(soa.elem, _I32), p(soa.ion, _I32), p(soa.wl), p(soa.gamma_rad), p(soa.gamma_stark),
        p(soa.gamma_vdw), p(soa.center_t, _I64), p(soa.center_w, _I64),
        p(_c(transp_arr, np.float64)), p(_c(valid, np.uint8), _U8), p(pop), p(dop),
        ct.c_long(pop.shape[1]), ct.c_long(pop.shape[2]), p(_xne(atm)),
        p(txnxn(atm)), p(_c(atm.temperature, np.float64)),
        p(_c(kapmin, np.float64)), ct.c_double(cutoff), ct.byref(t.struct), p(out), p(stats))
    return out, dict(e_line=stats[0], p_line=stats[1], wing_pairs=stats[2], sum_maxstep=stats[3],
                     max_maxstep=stats[4])


def lte_spectrum(grid, atm, cont_abs, cont_scat, line_abs, line_scat, line_source):
    tables()
    D, W = atm.layers, grid.size
    ft, fc = np.zeros(W), np.zeros(W)
    iters = np.zeros(1)
    iw = np.zeros(W, dtype=np.int32)
    args = [_c(a, np.float64) for a in (cont_abs, cont_scat, line_abs, line_scat)]
    ls = _c(line_source, np.float64) if line_source is not None else None
    lib().orc_lte_spectrum(ct.c_long(D), ct.c_long(W), p(_c(grid, np.float64)),
                           p(_c(atm.temperature, np.float64)), p(_c(atm.depth, np.float64)),
                           *[p(a) for a in args], p(ls) if ls is not None else None,
                           p(ft), p(fc), p(iters), p(iw, _I32))
    lte_spectrum.last_iters = iw
    return ft, fc, float(iters[0])


def lte_frequency(w, grid, atm, cont_abs, cont_scat, line_abs, line_scat, line_source):
    tables()
    D, W = atm.layers, grid.size
    fc = ct.c_double(0.0)
    info = np.zeros(3, dtype=np.intc)
    args = [_c(a, np.float64) for a in (cont_abs, cont_scat, line_abs, line_scat)]
    ls = _c(line_source, np.float64) if line_source is not None else None
    ft = lib().orc_lte_frequency(ct.c_long(D), ct.c_long(W), ct.c_long(w), ct.c_double(grid[w]),
                                 p(_c(atm.temperature, np.float64)), p(_c(atm.depth, np.float64)),
                                 *[p(a) for a in args], p(ls) if ls is not None else None,
                                 ct.byref(fc), p(info, _INT))
    return float(ft), float(fc.value), info.copy()


# ------------------------------------------------------------- hydrogen lines
HSTATE_FIELDS = ("pp", "fo", "y1b", "y1s", "t3nhe", "t3nh2", "c1d", "c2d", "gcon1", "gcon2", "xnfph0", "xnfph1",
                 "dopph", "xne", "has")


def vcse1f(x: float) -> float:
    tables()
    return lib().orc_vcse1f(ct.c_double(x))


def sofbeta(beta: float, p_: float, n: int, m: int) -> float:
    tables()
    return lib().orc_sofbeta(ct.c_double(beta), ct.c_double(p_), ct.c_int(n), ct.c_int(m))


def hf_nm(n: int, m: int) -> float:
    return lib().orc_hf_nm(ct.c_int(n), ct.c_int(m))


def fine_structure(n: int, m: int):
    tables()
    off, wt = np.zeros(40), np.zeros(40)
    k = lib().orc_fine_structure(ct.c_int(n), ct.c_int(m), p(off), p(wt))
    return off[:k].copy(), wt[:k].copy()


def hydrogen_states(atm, microturb_kms: float = 0.0) -> np.ndarray:
    """Per-depth hydrogen state [D, 15] (physics/populations.py:96-128, 210-231)."""
    D = atm.layers
    he1, h2 = getattr(atm, "xnf_he1", None), getattr(atm, "xnf_h2", None)
    xnfph, dopph = getattr(atm, "xnfph", None), getattr(atm, "dopph", None)
    vt = getattr(atm, "turbulent_velocity", None)
    opt = lambda a: None if a is None else _c(a, np.float64)  # noqa: E731
    out = np.zeros((D, len(HSTATE_FIELDS)))
    xnfph = opt(xnfph)
    if xnfph is not None:
        xnfph = np.ascontiguousarray(xnfph.reshape(D, -1)[:, :2])
    args = [opt(he1), opt(h2), xnfph, opt(dopph), opt(vt)]
    lib().orc_hydrogen_states(ct.c_long(D), p(_c(atm.temperature, np.float64)), p(_xne(atm)),
                              *[p(a) if a is not None else None for a in args], ct.c_double(microturb_kms),
                              ct.c_int(int(he1 is not None or h2 is not None)), p(out))
    return out


def hydrogen_profile(n: int, m: int, state: np.ndarray, delta_nm: float) -> float:
    tables()
    st = _c(state, np.float64)
    return lib().orc_hydrogen_profile(ct.c_int(n), ct.c_int(m), p(st), ct.c_double(delta_nm))


def hydrogen_line_selection(cat, grid):
    """The lines engine/opacity.py:1776-1782 takes (line type -1/-2, ion stage
    1), in catalog order: wavelength, centre index, gf, elo, n_lower, n_upper."""
    sel = [i for i, r in enumerate(cat.records) if int(r.line_type or 0) in (-1, -2) and r.ion_stage == 1]
    idx = np.array(sel, dtype=np.int64)
    centers = nearest_grid_indices(grid, np.ascontiguousarray(cat.index_wavelength, np.float64))
    return dict(index=idx, wl=_c(cat.wavelength[idx], np.float64), center=_c(centers[idx], np.int64),
                gf=_c(cat.gf[idx], np.float64), elo=_c(cat.excitation[idx], np.float64),
                n_lower=np.array([cat.records[i].n_lower for i in sel], np.int32),
                n_upper=np.array([cat.records[i].n_upper for i in sel], np.int32))


def hydrogen_lines(cat, atm, grid, cont, stim, cutoff=1e-3, microturb_kms=0.0):
    """AHLINE [D, W] (engine/opacity.py:1706-1884) and the profile evaluation count."""
    t = tables()
    D, W = atm.layers, grid.size
    h = hydrogen_line_selection(cat, grid)
    out = np.zeros((D, W))
    stats = np.zeros(1)
    if h["index"].size == 0 or atm.population_per_ion is None or atm.doppler_per_ion is None:
        return out, 0.0
    pop = np.ascontiguousarray(atm.population_per_ion[:, 0, 0], np.float64)
    dop = np.ascontiguousarray(atm.doppler_per_ion[:, 0, 0], np.float64)
    rho = getattr(atm, "mass_density", None)
    if rho is None:
        return out, 0.0
    st = hydrogen_states(atm, microturb_kms)
    rc = lib().orc_hydrogen_lines(
        ct.c_long(h["index"].size), ct.c_long(D), ct.c_long(W), p(_c(grid, np.float64)), p(h["wl"]),
        p(h["center"], _I64), p(h["gf"]), p(h["elo"]), p(h["n_lower"], _I32), p(h["n_upper"], _I32), p(pop),
        p(dop), p(_c(rho, np.float64)), p(hckt(atm)), p(_xne(atm)), p(st), p(_c(cont, np.float64)),
        p(_c(stim, np.float64)), ct.c_double(cutoff), ct.c_double(microturb_kms), p(out), p(stats))
    if rc != 0:
        raise ValueError("hydrogen line beyond the HPROF4 damping tables (n_upper + 2 > 96)")
    return out, float(stats[0])


# ------------------------------------------------------------- fort.19 lines
class _F19Tables(ct.Structure):
    _fields_ = [("h0", _D), ("h1", _D), ("h2", _D), ("nvoigt", ct.c_int),
                ("extab", _D), ("extabf", _D), ("nfastex", ct.c_int)]


def doppler_scale(atm, microturb_kms: float = 0.0) -> np.ndarray:
    """Doppler width / wavelength per depth for the engine's default line mass,
    56 amu (physics/populations.py:131-190; the engine passes no elements)."""
    out = []
    vt = getattr(atm, "turbulent_velocity", None)
    for i, t in enumerate(np.asarray(atm.temperature, np.float64).tolist()):
        t = max(t, 1.0)
        th = np.sqrt(2.0 * 1.380649e-16 * t / (np.array([56.0]) * 1.66054e-24)) / 2.99792458e10
        v = (float(vt[i]) if vt is not None and len(vt) > i else 0.0) / 1e5 / 299792.458
        m = microturb_kms / 299792.458 if microturb_kms > 0 else 0.0
        turb = math.sqrt(v * v + m * m)
        out.append(float(np.sqrt(turb ** 2 + th ** 2)[0]))
    return np.array(out, np.float64)


def fort19_selection(cat, f19, grid, n_ion, n_elem):
    """The records _add_fort19_asynth takes (engine/opacity.py:1364-1397), with
    their matched catalog index and centre."""
    import catalog as C_

    fw = np.array([r.wavelength for r in f19], np.float64)
    to_cat = {v: k for k, v in C_.match_fort19(np.asarray(cat.wavelength, np.float64), fw).items()}
    centers = nearest_grid_indices(grid, fw) if fw.size else np.zeros(0, np.int64)
    sel = []
    for f, r in enumerate(f19):
        kind = 0 if r.line_type == 0 else 1 if r.line_type == 1 else 2 if r.line_type > 3 else -1
        ci = to_cat.get(f)
        if kind < 0 or ci is None:
            continue
        rec = cat.records[ci]
        z = C_._ENGINE_Z.get(str(rec.element).strip().upper().replace(" ", ""))
        if z is None or rec.ion_stage <= 0:
            continue
        if int(centers[f]) >= grid.size:
            continue
        if rec.ion_stage - 1 >= n_ion or z - 1 >= n_elem:
            raise IndexError("fort.19 record outside the population table")
        sel.append((f, kind, ci, z - 1, rec.ion_stage, int(centers[f])))
    return sel


def fort19_asynth(cat, atm, grid, cont, stim, asynth_arr, wl_start, wl_end, resolution, cutoff=1e-3,
                  microturb_kms=0.0):
    """_add_fort19_asynth (engine/opacity.py:1335-1546) into asynth_arr [D, W],
    in place; returns the number of records taken."""
    import catalog as C_

    t = tables()
    if atm.population_per_ion is None:
        return 0
    f19 = C_.fort19_build(cat, wl_start, wl_end, resolution)
    pop = _c(atm.population_per_ion, np.float64)
    sel = fort19_selection(cat, f19, grid, pop.shape[1], pop.shape[2])
    if not sel:
        return 0
    D, W, N = atm.layers, grid.size, len(sel)
    hk = hckt(atm)
    vs = doppler_scale(atm, microturb_kms)
    boltz = np.zeros((N, D))
    dopp = np.zeros((N, D))
    cols = {k: [] for k in ("kind", "wl", "center", "gf", "gr", "gs", "gw", "elem", "ion", "nelion", "nlast",
                            "ncon", "nelionx")}
    for j, (f, kind, ci, el, st, c) in enumerate(sel):
        r = f19[f]
        e = float(cat.excitation[ci])
        for d in range(D):
            boltz[j, d] = lib().orc_fastex(ct.c_double(e * hk[d]), p(t.extab), p(t.extabf), ct.c_int(1001))
            dopp[j, d] = float(cat.wavelength[ci]) * vs[d]
        for k, v in (("kind", kind), ("wl", r.wavelength), ("center", c), ("gf", r.gf), ("gr", r.gamma_rad),
                     ("gs", r.gamma_stark), ("gw", r.gamma_vdw), ("elem", el), ("ion", st), ("nelion", r.nelion),
                     ("nlast", r.line_type), ("ncon", r.ncon), ("nelionx", r.nelionx)):
            cols[k].append(v)
    ne = np.maximum(_xne(atm), 1e-40)
    nmerge = np.maximum(1600.0 / np.power(ne, 2.0 / 15.0) - 1.5, 1.0)
    emerge = 109737.312 / np.maximum(nmerge ** 2, 1e-12)
    emerge_h = 109677.576 / np.maximum(nmerge ** 2, 1e-12)
    rho = getattr(atm, "mass_density", None)
    rho = np.zeros(D) if rho is None else _c(rho, np.float64)
    depth = dict(pop=pop, rho=rho, xne=_xne(atm), txnxn=txnxn(atm), emerge=emerge, emerge_h=emerge_h,
                 boltz=boltz, doppler=dopp)
    fort19_raw(grid, cols, depth, cont, stim, asynth_arr, cutoff)
    return N


def fort19_raw(grid, rec: dict, depth: dict, cont, stim, asynth_arr, cutoff=1e-3):
    """orc_fort19_asynth on explicit arrays: rec holds the per-record columns
    (kind, wl, center, gf, gr, gs, gw, elem, ion, nelion, nlast, ncon, nelionx),
    depth the per-depth ones (pop [D][nion][nelem], rho, xne, txnxn, emerge,
    emerge_h) and boltz, doppler [N][D]; asynth_arr [D, W] is added to in place."""
    t = tables()
    i32 = {k: np.ascontiguousarray(rec[k], np.int32) for k in ("kind", "elem", "ion", "nelion", "nlast", "ncon",
                                                               "nelionx")}
    f8 = {k: np.ascontiguousarray(rec[k], np.float64) for k in ("wl", "gf", "gr", "gs", "gw")}
    center = np.ascontiguousarray(rec["center"], np.int64)
    N = center.size
    pop = _c(depth["pop"], np.float64)
    D, W = pop.shape[0], grid.size
    with np.load(_HERE.parent / "kurucz_py_joss_submission_amd" / "data" / "kurucz_tables.npz",
                 allow_pickle=False) as z:
        contx = np.ascontiguousarray(z["contx"], np.float64)
    tb = _F19Tables(p(t.h0), p(t.h1), p(t.h2), 2001, p(t.extab), p(t.extabf), 1001)
    out = _c(asynth_arr, np.float64)
    dd = {k: _c(depth[k], np.float64) for k in ("rho", "xne", "txnxn", "emerge", "emerge_h", "boltz", "doppler")}
    lib().orc_fort19_asynth(
        ct.c_long(N), ct.c_long(D), ct.c_long(W), p(_c(grid, np.float64)), p(i32["kind"], _I32), p(f8["wl"]),
        p(center, _I64), p(f8["gf"]), p(f8["gr"]), p(f8["gs"]), p(f8["gw"]), p(i32["elem"], _I32),
        p(i32["ion"], _I32), p(i32["nelion"], _I32), p(i32["nlast"], _I32), p(i32["ncon"], _I32),
        p(i32["nelionx"], _I32), p(contx), p(dd["boltz"]), p(dd["doppler"]), p(pop), ct.c_long(pop.shape[1]),
        ct.c_long(pop.shape[2]), p(dd["rho"]), p(dd["xne"]), p(dd["txnxn"]), p(dd["emerge"]), p(dd["emerge_h"]),
        p(_c(cont, np.float64)), p(_c(stim, np.float64)), ct.c_double(cutoff), ct.byref(tb), p(out))
    if out is not asynth_arr:
        asynth_arr[...] = out
    if out is not asynth_arr:
        asynth_arr[...] = out
    return N


def helium_asynth(cat, atm, grid, cont, stim, out, wl_start, wl_end, resolution, cutoff=1e-3):
    """The engine's helium pass (engine/opacity.py:2534-2546, 2686-3066,
    3651-3659): each He line's Voigt wings, summed per depth, times STIM,
    added to out [D, W] in place.  Returns the number of lines taken."""
    import catalog as C_

    t = tables()
    pop, dop = getattr(atm, "population_per_ion", None), getattr(atm, "doppler_per_ion", None)
    if pop is None or dop is None:
        return 0
    pop, dop = np.asarray(pop, np.float64), np.asarray(dop, np.float64)
    f19 = C_.fort19_build(cat, wl_start, wl_end, resolution)
    match = C_.match_fort19(np.asarray(cat.wavelength, np.float64), np.array([r.wavelength for r in f19]))
    centers = nearest_grid_indices(grid, np.ascontiguousarray(cat.wavelength, np.float64))
    D, W = atm.layers, grid.size
    hk = hckt(atm)
    cols = {k: [] for k in ("wl", "center", "cgf", "gr", "gs", "gw", "ltype", "ion", "ncon", "nelionx")}
    popn, dopn, bz = [], [], []
    for i, rec in enumerate(cat.records):
        m = match.get(i)
        lt = f19[m].line_type if m is not None else int(rec.line_type)
        if lt not in (-3, -4, -6):
            continue
        sym = str(rec.element).strip().upper()
        if sym in {"H", "H I", "HI"} and rec.ion_stage == 1:
            continue
        c = int(centers[i])
        if c < 0 or c >= W:
            continue
        z = C_._ENGINE_Z.get(sym.replace(" ", ""))
        if z is None or z - 1 >= pop.shape[2] or rec.ion_stage > pop.shape[1]:
            continue
        lw = float(cat.wavelength[i])
        e = float(cat.excitation[i])
        for k, v in (("wl", lw), ("center", c), ("cgf", C_.CGF_FACTOR * float(cat.gf[i]) / (C_.C_NM / lw)),
                     ("gr", float(cat.gamma_rad[i])), ("gs", float(cat.gamma_stark[i])),
                     ("gw", float(cat.gamma_vdw[i])), ("ltype", lt), ("ion", rec.ion_stage),
                     ("ncon", f19[m].ncon if m is not None else 0),
                     ("nelionx", f19[m].nelionx if m is not None else 0)):
            cols[k].append(v)
        popn.append(pop[:, rec.ion_stage - 1, z - 1])
        dopn.append(dop[:, rec.ion_stage - 1, z - 1])
        bz.append([lib().orc_fastex(ct.c_double(e * hk[d]), p(t.extab), p(t.extabf), ct.c_int(1001))
                   for d in range(D)])
    N = len(cols["wl"])
    if N == 0:
        return 0
    f8 = {k: np.ascontiguousarray(cols[k], np.float64) for k in ("wl", "cgf", "gr", "gs", "gw")}
    i32 = {k: np.ascontiguousarray(cols[k], np.int32) for k in ("ltype", "ion", "ncon", "nelionx")}
    center = np.ascontiguousarray(cols["center"], np.int64)
    ne = np.maximum(_xne(atm), 1e-40)
    nmerge = np.maximum(1600.0 / np.power(ne, 2.0 / 15.0) - 1.5, 1.0)
    emerge = 109737.312 / np.maximum(nmerge ** 2, 1e-12)
    emerge_h = 109677.576 / np.maximum(nmerge ** 2, 1e-12)
    rho = getattr(atm, "mass_density", None)
    rho = np.ones(D) if rho is None else _c(rho, np.float64)
    with np.load(_HERE.parent / "kurucz_py_joss_submission_amd" / "data" / "kurucz_tables.npz",
                 allow_pickle=False) as z:
        contx = np.ascontiguousarray(z["contx"], np.float64)
    tb = _F19Tables(p(t.h0), p(t.h1), p(t.h2), 2001, p(t.extab), p(t.extabf), 1001)
    res = _c(out, np.float64)
    lib().orc_helium_wings(
        ct.c_long(N), ct.c_long(D), ct.c_long(W), p(_c(grid, np.float64)), p(f8["wl"]), p(center, _I64),
        p(f8["cgf"]), p(f8["gr"]), p(f8["gs"]), p(f8["gw"]), p(i32["ltype"], _I32), p(i32["ion"], _I32),
        p(i32["ncon"], _I32), p(i32["nelionx"], _I32), p(contx), p(_c(np.array(popn), np.float64)),
        p(_c(np.array(dopn), np.float64)), p(_c(np.array(bz), np.float64)), p(rho), p(_xne(atm)), p(txnxn(atm)),
        p(emerge), p(emerge_h), p(_c(cont, np.float64)), p(_c(stim, np.float64)), ct.c_double(cutoff),
        ct.byref(tb), p(res))
    if res is not out:
        out[...] = res
    return N


@dataclass
class OracleResult:
    wavelength: np.ndarray
    flux: np.ndarray  # per nm
    continuum: np.ndarray  # per nm
    flux_hz: np.ndarray
    cont_hz: np.ndarray
    stages: Dict[str, np.ndarray] = field(default_factory=dict)
    stats: Dict[str, float] = field(default_factory=dict)
    timings: Dict[str, float] = field(default_factory=dict)


def synthesize(atm, lines, wl_start, wl_end, resolution, cutoff=1e-3, microturb_kms=0.0,
               rhoxj=0.0, keep_stages=True, threads: Optional[int] = None,
               helium_wings: bool = True) -> OracleResult:
    """Full hot path on the CPU (engine/opacity.py:2084-3861 stage order).

    lines: the oracle's own catalog (catalog.Catalog, from catalog.read_gfall)
    or the path of a GFALL file, which is read and DELLIM-filtered to the
    window here.  The line arrays come from the oracle's restatement of the
    reference's prep (catalog.line_soa), not from the package under test."""
    import time

    import catalog as C_

    if isinstance(lines, (str, os.PathLike)):
        lines = C_.filter_by_range(C_.read_gfall(lines), wl_start, wl_end)
    if not isinstance(lines, C_.Catalog):
        raise TypeError("oracle.synthesize takes the oracle's own catalog (catalog.read_gfall) or a GFALL path")
    if threads:
        set_threads(threads)
    tm = {}
    t0 = time.perf_counter()
    grid = wavelength_grid(wl_start, wl_end, resolution)
    soa = C_.line_soa(lines, grid, atm.population_per_ion.shape[2], grid_origin=float(grid[0]))
    tm["prep"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    ca, cs = continuum(atm, grid)
    kapmin = ca + cs
    ehvkt, stim, bnu, slinec = freq_quantities(grid, atm.temperature)
    tm["continuum"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    tr, valid = transp(soa, atm, grid, kapmin, cutoff, microturb_kms)
    tm["transp"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    asy, st = asynth(soa, atm, grid, tr, valid, kapmin, cutoff)
    # engine/opacity.py:2366-2387: the fort.19 records join ASYNTH before the hydrogen lines
    st["fort19_records"] = float(fort19_asynth(lines, atm, grid, kapmin, stim, asy, wl_start, wl_end, resolution,
                                               cutoff, microturb_kms))
    ahl, st["h_profile_evals"] = hydrogen_lines(lines, atm, grid, kapmin, stim, cutoff, microturb_kms)
    asy = asy + ahl  # engine/opacity.py:2483-2487: AHLINE joins ASYNTH before the FSCAT split
    # engine/opacity.py:3551-3590, 3651-3659: the helium pass, HELIUM_WINGS * STIM, after the hydrogen lines
    st["helium_lines"] = float(helium_asynth(lines, atm, grid, kapmin, stim, asy, wl_start, wl_end, resolution,
                                             cutoff) if helium_wings else 0)
    tm["asynth"] = time.perf_counter() - t0
    rhox = atm.depth
    fscat = np.exp(-rhox / rhoxj) if rhoxj > 0.0 else np.zeros_like(rhox)
    aline = asy * (1.0 - fscat[:, None])
    sigl = asy * fscat[:, None]
    t0 = time.perf_counter()
    ft, fc, iters = lte_spectrum(grid, atm, ca, cs, aline, sigl, slinec)
    tm["rt"] = time.perf_counter() - t0
    st["josh_iters"] = iters
    st["valid_pairs"] = float(valid.sum())
    flux = ft.copy()
    cont = fc.copy()
    lib().orc_flux_to_lambda(ct.c_long(grid.size), p(grid), p(flux))
    lib().orc_flux_to_lambda(ct.c_long(grid.size), p(grid), p(cont))
    stages = {}
    if keep_stages:
        stages = dict(cont_abs=ca, cont_scat=cs, transp=tr, valid=valid, asynth=asy, ahline=ahl, slinec=slinec,
                      center_t=soa.center_t, center_w=soa.center_w, process=soa.process,
                      iters=lte_spectrum.last_iters)
    return OracleResult(grid, flux, cont, ft, fc, stages, st, tm)


def wing_params(soa, atm, grid, transp_arr, valid, kapmin, cutoff=1e-3):
    """Per-(line, depth) wing parameters: (doubles [L,D,6], ints [L,D,3]).

    doubles: kappa0, adamp, dvoigt, x_far, kapmin_ref, dopple;
    ints: n10dop, maxstep, use_far (-1 where the pair has no wing setup)."""
    t = tables()
    L, D, W = len(soa), atm.layers, grid.size
    pop = _c(atm.population_per_ion, np.float64)
    dop = _c(atm.doppler_per_ion, np.float64)
    od = np.zeros((L, D, 6))
    oi = np.zeros((L, D, 3), dtype=np.int64)
    lib().orc_wing_params(
        ct.c_long(L), ct.c_long(D), ct.c_long(W), p(_c(grid, np.float64)), p(soa.ltype, _I32),
        p(soa.elem, _I32), p(soa.ion, _I32), p(soa.wl), p(soa.gamma_rad), p(soa.gamma_stark),
        p(soa.gamma_vdw), p(soa.center_w, _I64), p(_c(transp_arr, np.float64)),
        p(_c(valid, np.uint8), _U8), p(pop), p(dop), ct.c_long(pop.shape[1]), ct.c_long(pop.shape[2]),
        p(_xne(atm)), p(txnxn(atm)),
        p(_c(kapmin, np.float64)), ct.c_double(cutoff), ct.byref(t.struct), p(od), p(oi, _I64))
    return od, oi
