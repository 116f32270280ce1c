"""Drop-in replacements for the reference's hot-path physics functions.

Same names, argument meaning, return shapes and error behaviour as the
reference (paths relative to its synthe_py/ package); the compute runs in
libsynthe_hip on the GPU.  Inputs and outputs are numpy arrays, so callers of
the reference can switch module imports without other changes:

* compute_transp                 physics/line_opacity.py:620-860
* compute_asynth_from_transp     physics/line_opacity.py:946-1340
* build_depth_continuum          physics/continuum.py:73-303
* solve_josh_flux                physics/josh_solver.py:367-767
* solve_lte_frequency            engine/radiative.py:42-218
* solve_lte_spectrum             engine/radiative.py:282-494
* compute_bfudge_and_slinec      physics/bfudge.py:72-168

The Boltzmann factors and TXNXN that the reference takes from a Populations
object (physics/populations.py:131-240) are computed on the device from the
atmosphere's HCKT and TXNXN (or its xnf_h/xnf_he1/xnf_h2 fallback), so the
``populations`` arguments are accepted and ignored.
"""

from __future__ import annotations

from typing import Optional, Tuple

import numpy as np

from . import _native as N
from .atmosphere import AtmosphereModel
from .lines import LineCatalog, build_line_soa, nearest_grid_indices


def _t():
    import torch

    return torch


def _dev(a, dtype=np.float64):
    from .device import to_device

    return to_device(np.asarray(a, dtype=dtype), dtype)


def _host(t) -> np.ndarray:
    return t.cpu().numpy()


def build_depth_continuum(atmosphere: AtmosphereModel, wavelength_grid: np.ndarray):
    """physics/continuum.py:73-303 -> (absorption[D,W], scattering[D,W], None, None)."""
    from .device import upload_atmosphere

    torch = _t()
    wl = np.asarray(wavelength_grid, dtype=np.float64)
    d = upload_atmosphere(atmosphere)
    D, W = atmosphere.layers, wl.size
    ca = torch.empty((D, W), dtype=torch.float64, device="cuda")
    cs = torch.empty((D, W), dtype=torch.float64, device="cuda")
    ctx = N.context()
    d_wl = _dev(wl)  # device tensors stay referenced until the launch is enqueued
    N.check("synthe_continuum", ctx.lib.synthe_continuum(
        ctx.handle, D, W, N.ptr(d_wl), int(d.wledge.shape[0]), N.ptr(d.wledge),
        N.ptr(d.half_edge), N.ptr(d.delta_edge), N.ptr(d.coeff_abs), N.ptr(d.coeff_scat),
        int(d.coeff_log10), N.ptr(ca), N.ptr(cs), N.stream_handle()))
    return _host(ca), _host(cs), None, None


def compute_transp(catalog: LineCatalog, populations, atmosphere: AtmosphereModel,
                   cutoff: float = 1e-3, continuum_absorption: Optional[np.ndarray] = None,
                   wavelength_grid: Optional[np.ndarray] = None,
                   continuum_absorption_full: Optional[np.ndarray] = None,
                   wavelength_grid_full: Optional[np.ndarray] = None,
                   microturb_kms: float = 0.0) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    """physics/line_opacity.py:620-860 -> (transp[L,D], valid_mask[L,D], center_indices[L])."""
    if continuum_absorption is None or wavelength_grid is None:
        raise ValueError("continuum_absorption and wavelength_grid are required for compute_transp. "
                         "Fortran always uses KAPMIN = CONTINUUM * CUTOFF with no fallback.")
    if continuum_absorption_full is not None and wavelength_grid_full is not None:
        raise NotImplementedError("full-grid KAPMIN (continuum_absorption_full) is not on the hot path")
    from .device import upload_atmosphere, upload_lines

    torch = _t()
    wl = np.asarray(wavelength_grid, dtype=np.float64)
    d = upload_atmosphere(atmosphere)
    soa = build_line_soa(catalog, wl, d.n_elem)
    lines = upload_lines(soa, wl.size)
    L, D = len(soa), atmosphere.layers
    tr = torch.zeros((L, D), dtype=torch.float64, device="cuda")
    va = torch.zeros((L, D), dtype=torch.uint8, device="cuda")
    if L:
        ctx = N.context()
        ls, ds = lines.struct(), d.struct()
        import ctypes

        d_km = _dev(continuum_absorption)
        N.check("synthe_transp", ctx.lib.synthe_transp(
            ctx.handle, ctypes.byref(ls), ctypes.byref(ds), wl.size, N.ptr(d_km), float(cutoff),
            float(microturb_kms), N.ptr(tr), N.ptr(va), N.stream_handle()))
    return _host(tr), _host(va).astype(bool), soa.center_t.copy()


def compute_asynth_from_transp(transp: np.ndarray, catalog: LineCatalog, atmosphere: AtmosphereModel,
                               wavelength_grid: np.ndarray, valid_mask: Optional[np.ndarray] = None,
                               populations=None, cutoff: float = 1e-3,
                               continuum_absorption: Optional[np.ndarray] = None,
                               continuum_absorption_full: Optional[np.ndarray] = None,
                               wavelength_grid_full: Optional[np.ndarray] = None,
                               metal_tables=None, grid_origin: Optional[float] = None) -> np.ndarray:
    """physics/line_opacity.py:946-1340 -> asynth[D,W] (stimulated emission applied).

    Without continuum_absorption the cutoff is off, as in the reference: near
    wings run to N10DOP and far wings are skipped (KAPMIN_REF = 0)."""
    if continuum_absorption_full is not None and wavelength_grid_full is not None:
        raise NotImplementedError("full-grid KAPMIN (continuum_absorption_full) is not on the hot path")
    if metal_tables is not None:
        raise NotImplementedError("WCON/WTAIL merged-continuum tapers (fort.19 lines) are out of scope")
    from .device import check_wing_reach, upload_atmosphere, upload_lines

    torch = _t()
    wl = np.asarray(wavelength_grid, dtype=np.float64)
    D, W = atmosphere.layers, wl.size
    check_wing_reach(wl, getattr(atmosphere, "doppler_per_ion", None))
    d = upload_atmosphere(atmosphere)
    soa = build_line_soa(catalog, wl, d.n_elem, grid_origin=grid_origin)
    lines = upload_lines(soa, W)
    L = len(soa)
    tr = np.asarray(transp, dtype=np.float64)
    if tr.shape != (L, D):
        raise ValueError(f"transp has shape {tr.shape}, expected {(L, D)}")
    va = np.ones((L, D), dtype=np.uint8) if valid_mask is None else np.asarray(valid_mask, dtype=np.uint8)
    use_cutoff = continuum_absorption is not None and np.shape(continuum_absorption) == (D, W)
    kapmin = np.asarray(continuum_absorption, np.float64) if use_cutoff else np.zeros((D, W))
    out = torch.empty((D, W), dtype=torch.float64, device="cuda")
    import ctypes

    ctx = N.context()
    ls, ds = lines.struct(), d.struct()
    d_wl, d_tr, d_va, d_km = _dev(wl), _dev(tr), _dev(va, np.uint8), _dev(kapmin)
    N.check("synthe_asynth", ctx.lib.synthe_asynth(
        ctx.handle, ctypes.byref(ls), ctypes.byref(ds), W, N.ptr(d_wl), N.ptr(d_tr), N.ptr(d_va),
        N.ptr(d_km), float(cutoff) if use_cutoff else 0.0, N.ptr(lines.corder),
        N.ptr(lines.cstart), N.ptr(lines.worder), N.ptr(out), None, N.stream_handle()))
    return _host(out)


def compute_hydrogen_line_opacity(catalog, pops, atmosphere_model: AtmosphereModel, wavelength_grid: np.ndarray,
                                  continuum: np.ndarray, stim: np.ndarray, cutoff: float,
                                  microturb_kms: float = 0.0) -> np.ndarray:
    """engine/opacity.py:1706-1884 (_compute_hydrogen_line_opacity) -> AHLINE [D, W].

    The catalog's H I lines (type -1/-2, ion stage 1) with their HPROF4
    profiles, walked as engine/opacity.py:1547-1704 walks them.  The per-depth
    quantities the reference takes from pops (Boltzmann factors, the hydrogen
    state) are computed from atmosphere_model exactly as
    physics/populations.py:130-240 computes them, so pops is not read; stim and
    continuum [D, W] are used as given.  Zero where the reference returns
    zero (no populations, Doppler widths or mass density)."""
    from .atmosphere import depth_hckt
    from .lines import hydrogen_lines

    torch = _t()
    atm = atmosphere_model
    wl = np.asarray(wavelength_grid, dtype=np.float64)
    D, W = atm.layers, wl.size
    out = torch.zeros((D, W), dtype=torch.float64, device="cuda")
    h = hydrogen_lines(catalog, wl)
    pop, dop, rho = (getattr(atm, n, None) for n in ("population_per_ion", "doppler_per_ion", "mass_density"))
    if len(h) == 0 or pop is None or dop is None or rho is None or np.asarray(pop).shape[2] < 1:
        return _host(out)
    f = lambda v: None if v is None else np.ascontiguousarray(v, np.float64)  # noqa: E731
    xnfph = getattr(atm, "xnfph", None)
    keep = [f(atm.temperature), f(atm.electron_density), f(rho), f(depth_hckt(atm)),
            f(np.asarray(pop)[:, 0, 0]), f(np.asarray(dop)[:, 0, 0]), f(getattr(atm, "xnf_he1", None)),
            f(getattr(atm, "xnf_h2", None)),
            None if xnfph is None else np.ascontiguousarray(np.asarray(xnfph, np.float64).reshape(D, -1)[:, :2]),
            f(getattr(atm, "dopph", None)), f(getattr(atm, "turbulent_velocity", None))]
    hl = N.HLines(len(h), *(x.ctypes.data for x in (h.wavelength, h.center, h.gf, h.elo, h.n_lower, h.n_upper)))
    hd = N.HDepths(D, *(None if x is None else x.ctypes.data for x in keep))
    cont = np.asarray(continuum, np.float64)
    st = np.asarray(stim, np.float64)
    if cont.shape != (D, W) or st.shape != (D, W):
        raise ValueError(f"continuum and stim must be [D, W] = {(D, W)}")
    import ctypes

    ctx = N.context()
    d_wl, d_c, d_s = _dev(wl), _dev(cont), _dev(st)
    N.check("synthe_hydrogen_lines", ctx.lib.synthe_hydrogen_lines(
        ctx.handle, ctypes.byref(hl), ctypes.byref(hd), float(microturb_kms), float(cutoff), W, N.ptr(d_wl),
        N.ptr(d_c), N.ptr(d_s), N.ptr(out), N.stream_handle()))
    return _host(out)


def add_fort19_asynth(asynth: np.ndarray, stim: np.ndarray, wavelength: np.ndarray, continuum: np.ndarray,
                      contx: np.ndarray, emerge: np.ndarray, emerge_h: np.ndarray, catalog, fort19_data,
                      catalog_to_fort19, pops, atm: AtmosphereModel, cutoff: float) -> None:
    """engine/opacity.py:1335-1546 (_add_fort19_asynth): the fort.19 records'
    terms (normal lines with NBLO/NBUP, autoionising Shore profiles, merged-
    continuum ramps) times STIM, added into asynth [D, W] in place.  The
    per-line Boltzmann factors, Doppler widths, electron densities and TXNXN
    are taken from pops as the reference takes them (computed from atm when
    pops is None); continuum, stim, emerge and contx are used as given."""
    from .atmosphere import depth_doppler_scale, depth_hckt, depth_txnxn
    from .lines import fort19_lines

    if fort19_data is None or len(fort19_data.wavelength_vacuum) == 0:
        return
    if atm.population_per_ion is None:
        return
    wl = np.asarray(wavelength, dtype=np.float64)
    pop = np.ascontiguousarray(atm.population_per_ion, np.float64)
    D, W = atm.layers, wl.size
    f = fort19_lines(catalog, fort19_data, wl, pop.shape[1], pop.shape[2], catalog_to_fort19, contx)
    if len(f) == 0:
        return
    c = np.ascontiguousarray
    rho = np.zeros(D) if atm.mass_density is None else c(atm.mass_density, np.float64)
    if pops is not None:
        layers = [pops.layers[d] for d in range(D)]
        boltz = c(np.array([[float(layers[d].boltzmann_factor[i]) for d in range(D)] for i in f.cat_index]).reshape(
            len(f), D))
        dopp = c(np.array([[float(layers[d].doppler_width[i]) for d in range(D)] for i in f.cat_index]).reshape(
            len(f), D))
        xne = c([float(layers[d].electron_density) for d in range(D)], np.float64)
        tx = c([float(layers[d].txnxn) for d in range(D)], np.float64)
        hckt = vscale = None
    else:
        boltz = dopp = None
        xne = c(atm.electron_density, np.float64)
        tx = c(depth_txnxn(atm))
        hckt, vscale = c(depth_hckt(atm)), c(depth_doppler_scale(atm))
    keep = [pop, rho, xne, tx, c(atm.temperature, np.float64), c(emerge, np.float64), c(emerge_h, np.float64),
            boltz, hckt, dopp, vscale]
    fl = N.F19Lines(len(f), *(x.ctypes.data for x in (f.kind, f.wavelength, f.center, f.gf, f.gamma_rad,
                                                     f.gamma_stark, f.gamma_vdw, f.elem, f.ion, f.nelion, f.nlast,
                                                     f.cont_val, f.cat_wavelength, f.elo)))
    fd = N.F19Depths(D, pop.shape[1], pop.shape[2], *(None if x is None else x.ctypes.data for x in keep))
    cont = np.asarray(continuum, np.float64)
    st = np.asarray(stim, np.float64)
    if cont.shape != (D, W) or st.shape != (D, W) or np.shape(asynth) != (D, W):
        raise ValueError(f"asynth, continuum and stim must be [D, W] = {(D, W)}")
    import ctypes

    ctx = N.context()
    d_wl, d_c, d_s, d_a = _dev(wl), _dev(cont), _dev(st), _dev(asynth)
    N.check("synthe_fort19_lines", ctx.lib.synthe_fort19_lines(
        ctx.handle, ctypes.byref(fl), ctypes.byref(fd), float(cutoff), W, N.ptr(d_wl), N.ptr(d_c), N.ptr(d_s),
        N.ptr(d_a), N.stream_handle()))
    asynth[...] = _host(d_a)


def _departure_factor(data, layers: int) -> np.ndarray:
    """physics/bfudge.py:26-69 (_prep_factor with one column): a missing table
    is ones; a [D] or [D][k] table gives its first column; another row count
    is a ValueError."""
    if data is None:
        return np.ones(layers, dtype=np.float64)
    arr = np.asarray(data, dtype=np.float64)
    if arr.ndim == 1:
        arr = arr[:, None]
    if arr.ndim != 2:
        raise ValueError("Expected 2-D array for BFUDGE inputs")
    if arr.shape[0] != layers:
        raise ValueError(f"BFUDGE helper expected {layers} layers, received {arr.shape[0]}")
    if arr.shape[1] == 0:
        return np.ones(layers, dtype=np.float64)
    return arr[:, 0]


def bfudge_column(atmosphere: AtmosphereModel, params, layers: Optional[int] = None) -> np.ndarray:
    """BFUDGE [D] = BHYD^PH1 (BC1/BC2)^PC1 (BSI1/BSI2)^PSI1 (physics/bfudge.py:
    108-115, spectrv.for 173-174), each ratio and base floored at 1e-30;
    exactly 1 in LTE (PH1 = PC1 = PSI1 = 0).  params: a spectrv.SpectrvParams
    (None = LTE)."""
    n = int(layers if layers is not None else atmosphere.layers)
    ph1, pc1, psi1 = ((0.0, 0.0, 0.0) if params is None else
                      (float(params.ph1), float(params.pc1), float(params.psi1)))
    bhyd = _departure_factor(getattr(atmosphere, "bhyd", None), n)
    bc1 = _departure_factor(getattr(atmosphere, "bc1", None), n)
    bc2 = _departure_factor(getattr(atmosphere, "bc2", None), n)
    bsi1 = _departure_factor(getattr(atmosphere, "bsi1", None), n)
    bsi2 = _departure_factor(getattr(atmosphere, "bsi2", None), n)
    with np.errstate(over="ignore", under="ignore", divide="ignore", invalid="ignore"):
        return (np.power(np.maximum(bhyd, 1e-30), ph1)
                * np.power(np.maximum(bc1 / np.maximum(bc2, 1e-30), 1e-30), pc1)
                * np.power(np.maximum(bsi1 / np.maximum(bsi2, 1e-30), 1e-30), psi1))


def compute_bfudge_and_slinec(atmosphere: AtmosphereModel, params, bnu: np.ndarray, stim: np.ndarray,
                              ehvkt: np.ndarray, *, clip=None):
    """physics/bfudge.py:72-168: (BFUDGE [D], SLINEC [D][W] = BNU STIM / (BFUDGE -
    EHVKT)), elementwise on the caller's host arrays as the reference
    computes it (no clamping; NaN/inf propagate; clip is accepted and
    ignored, as there).  The device path computes SLINEC inside the source
    function kernel (synthe_source_function_bfudge) from the same BFUDGE."""
    layers = bnu.shape[0]
    bfudge = bfudge_column(atmosphere, params, layers)
    with np.errstate(over="ignore", under="ignore", divide="ignore", invalid="ignore"):
        slinec = (bnu * stim) / (bfudge[:, None] - ehvkt)
    return bfudge, slinec


def source_function(wavelength_grid: np.ndarray, temperature: np.ndarray):
    """engine/opacity.py:2215-2228 on the device -> (bnu, stim, ehvkt, slinec) [D,W]."""
    torch = _t()
    wl = np.asarray(wavelength_grid, np.float64)
    T = np.asarray(temperature, np.float64)
    D, W = T.size, wl.size
    outs = [torch.empty((D, W), dtype=torch.float64, device="cuda") for _ in range(4)]
    ctx = N.context()
    d_wl, d_t = _dev(wl), _dev(T)
    N.check("synthe_source_function", ctx.lib.synthe_source_function(
        ctx.handle, D, W, N.ptr(d_wl), N.ptr(d_t), *[N.ptr(o) for o in outs], N.stream_handle()))
    return tuple(_host(o) for o in outs)


def solve_lte_spectrum(wavelength_nm: np.ndarray, temperature: np.ndarray, column_mass: np.ndarray,
                       cont_abs: np.ndarray, cont_scat: np.ndarray, line_opacity: np.ndarray,
                       line_scattering: np.ndarray, line_source: Optional[np.ndarray] = None,
                       n_workers: Optional[int] = None) -> Tuple[np.ndarray, np.ndarray]:
    """engine/radiative.py:282-494 -> (flux_total[W], flux_cont[W]) per Hz.

    n_workers is accepted for signature compatibility; the GPU solves every
    wavelength in parallel."""
    torch = _t()
    wl = np.asarray(wavelength_nm, np.float64)
    W = wl.size
    D = int(np.asarray(temperature).size)
    ft = torch.empty(W, dtype=torch.float64, device="cuda")
    fc = torch.empty(W, dtype=torch.float64, device="cuda")
    if W == 0:
        return np.zeros(0), np.zeros(0)
    ctx = N.context()
    arrs = [_dev(a) for a in (cont_abs, cont_scat, line_opacity, line_scattering)]
    ls = _dev(line_source) if line_source is not None else None
    d_wl, d_t, d_m = _dev(wl), _dev(temperature), _dev(column_mass)
    N.check("synthe_lte_spectrum", ctx.lib.synthe_lte_spectrum(
        ctx.handle, D, W, N.ptr(d_wl), N.ptr(d_t), N.ptr(d_m), *[N.ptr(a) for a in arrs],
        N.ptr(ls), N.ptr(ft), N.ptr(fc), None, N.stream_handle()))
    return _host(ft), _host(fc)


def solve_lte_frequency(wavelength_nm: float, temperature: np.ndarray, column_mass: np.ndarray,
                        cont_abs: np.ndarray, cont_scat: np.ndarray, line_opacity: np.ndarray,
                        line_scattering: np.ndarray, line_source: Optional[np.ndarray] = None
                        ) -> Tuple[float, float]:
    """engine/radiative.py:42-218 for one wavelength -> (flux_total, flux_cont)."""
    col = lambda a: None if a is None else np.asarray(a, np.float64).reshape(-1, 1)  # noqa: E731
    ft, fc = solve_lte_spectrum(np.array([wavelength_nm], np.float64), temperature, column_mass,
                                col(cont_abs), col(cont_scat), col(line_opacity), col(line_scattering),
                                col(line_source))
    return float(ft[0]), float(fc[0])


def solve_josh_flux(acont: np.ndarray, scont: np.ndarray, aline: np.ndarray, sline: np.ndarray,
                    sigmac: np.ndarray, sigmal: np.ndarray, column_mass: np.ndarray) -> float:
    """physics/josh_solver.py:367-767 for one frequency -> HNU (surface flux)."""
    a = np.asarray(acont, np.float64)
    if a.size == 0:
        return 0.0
    return float(solve_josh_flux_batch(*(np.asarray(x, np.float64).reshape(-1, 1) for x in
                                         (acont, scont, aline, sline, sigmac, sigmal)),
                                       np.asarray(column_mass, np.float64))[0])


def solve_josh_flux_batch(acont, scont, aline, sline, sigmac, sigmal, column_mass) -> np.ndarray:
    """solve_josh_flux over columns: inputs [n, ncol], column_mass [n] -> flux[ncol]."""
    torch = _t()
    n, ncol = np.shape(acont)
    out = torch.empty(ncol, dtype=torch.float64, device="cuda")
    ctx = N.context()
    devs = [_dev(x) for x in (acont, scont, aline, sline, sigmac, sigmal, column_mass)]
    N.check("synthe_josh_flux_batch", ctx.lib.synthe_josh_flux_batch(
        ctx.handle, n, ncol, *[N.ptr(x) for x in devs], N.ptr(out), None, N.stream_handle()))
    return _host(out)


def nearest_grid_indices_ref(grid, values):
    """engine/opacity.py:550-592 (host)."""
    return nearest_grid_indices(np.asarray(grid, np.float64), np.asarray(values, np.float64))
