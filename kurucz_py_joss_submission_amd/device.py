"""Device-resident spectrum plans: inputs uploaded once, every stage on the GPU.

A SpectrumPlan holds one atmosphere, one line list and one wavelength grid in
HBM and runs the hot path (continuum -> source function -> TRANSP -> ASYNTH ->
JOSH RT -> unit conversion) as a chain of libsynthe_hip launches on the
current torch stream.  PyTorch is used only to own device memory and the stream.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import _native as N
from .atmosphere import (AtmosphereModel, depth_doppler_scale, depth_hckt, depth_rho, depth_txnxn, depth_xne,
                         merge_energies)
from .lines import (LineCatalog, LineSoA, build_line_soa, fort19_lines, fort19_records, helium_wing_lines,
                    hydrogen_lines)


def _torch():
    import torch

    return torch


def to_device(a: np.ndarray, dtype=None):
    torch = _torch()
    arr = np.ascontiguousarray(a if dtype is None else np.asarray(a, dtype=dtype))
    return torch.from_numpy(arr).to("cuda", non_blocking=False)


def check_wing_reach(grid: np.ndarray, doppler_per_ion) -> None:
    """Reject a grid so fine that a Voigt zone, N10DOP = 10 (v/c) R
    (physics/line_opacity.py:419-420), could pass MAXPROF = 10^6 pixels: the
    kernels bound every zone there, where the reference's near-zone loop
    would go on."""
    from .constants import MAX_PROFILE_STEPS

    g = np.asarray(grid, np.float64)
    if g.size < 2 or not g[1] / g[0] > 1.0 or doppler_per_ion is None:
        return
    resolu = 1.0 / (g[1] / g[0] - 1.0)
    dop = np.asarray(doppler_per_ion, np.float64)
    worst = 10.0 * float(np.max(dop, initial=0.0)) * resolu
    if worst > MAX_PROFILE_STEPS:
        raise ValueError(f"grid resolving power {resolu:.4g} with Doppler widths up to {np.max(dop):.3g} c "
                         f"gives Voigt zones of {worst:.4g} pixels, above MAXPROF = {MAX_PROFILE_STEPS}")


def center_csr(center_t: np.ndarray, n_wl: int):
    """Lines stably sorted by centre index and per-pixel offsets (host prep)."""
    order = np.argsort(center_t, kind="stable").astype(np.int64)
    ct_sorted = center_t[order]
    start = np.searchsorted(ct_sorted, np.arange(n_wl + 1), side="left").astype(np.int64)
    return order, start


@dataclass
class DepthArrays:
    temperature: object
    rhox: object
    rho: object
    xne: object
    txnxn: object
    hckt: object
    pop: object
    dop: object
    n_ion: int
    n_elem: int
    wledge: object
    half_edge: Optional[object]
    delta_edge: Optional[object]
    coeff_abs: object
    coeff_scat: object
    coeff_log10: bool

    @property
    def n(self) -> int:
        return int(self.temperature.shape[0])

    def struct(self) -> N.Depths:
        p = N.ptr
        return N.Depths(self.n, p(self.temperature), p(self.rhox), p(self.rho), p(self.xne),
                        p(self.txnxn), p(self.hckt), p(self.pop), p(self.dop), self.n_ion,
                        self.n_elem)


def upload_atmosphere(atm: AtmosphereModel) -> DepthArrays:
    """Device copies of a bundle's per-depth arrays.  atm is this package's
    AtmosphereModel or any object with the reference's fields
    (io/atmosphere.py:16-79); missing optional columns take the reference's
    fallbacks."""
    if getattr(atm, "population_per_ion", None) is None or getattr(atm, "doppler_per_ion", None) is None:
        raise ValueError("atmosphere bundle lacks population_per_ion / doppler_per_ion")
    if atm.population_per_ion is None or atm.doppler_per_ion is None:
        raise ValueError("atmosphere bundle lacks population_per_ion / doppler_per_ion")
    if atm.continuum_wledge is None or atm.continuum_abs_coeff is None or atm.continuum_scat_coeff is None:
        raise ValueError("Continuum coefficients are missing from atmosphere model "
                         "(need wledge, cont_abs_coeff, cont_scat_coeff)")
    f8 = np.float64
    wledge = np.asarray(atm.continuum_wledge, f8)
    E = wledge.size
    if E < 2:
        raise ValueError("Continuum edge table requires at least two wavelength entries")
    half = atm.continuum_half_edge
    delta = atm.continuum_delta_edge
    pop = np.asarray(atm.population_per_ion, f8)
    ca = np.asarray(atm.continuum_abs_coeff, f8)
    cs = np.asarray(atm.continuum_scat_coeff, f8)
    if ca.shape[1] != E - 1 or cs.shape[1] != E - 1:
        raise ValueError("Continuum coefficient grid does not match edge table dimensions")
    return DepthArrays(
        temperature=to_device(atm.temperature, f8),
        rhox=to_device(atm.depth, f8),
        rho=to_device(depth_rho(atm), f8),
        xne=to_device(depth_xne(atm), f8),
        txnxn=to_device(depth_txnxn(atm), f8),
        hckt=to_device(depth_hckt(atm), f8),
        pop=to_device(pop, f8),
        dop=to_device(atm.doppler_per_ion, f8),
        n_ion=int(pop.shape[1]),
        n_elem=int(pop.shape[2]),
        wledge=to_device(wledge, f8),
        half_edge=to_device(half, f8) if half is not None and np.size(half) == E - 1 else None,
        delta_edge=to_device(delta, f8) if delta is not None and np.size(delta) == E - 1 else None,
        coeff_abs=to_device(ca, f8),
        coeff_scat=to_device(cs, f8),
        coeff_log10=bool(atm.continuum_coeff_log10),
    )


@dataclass
class LineArrays:
    soa: LineSoA
    t: Dict[str, object]
    corder: object
    cstart: object
    worder: object

    @property
    def n(self) -> int:
        return len(self.soa)

    def struct(self) -> N.Lines:
        p = N.ptr
        t = self.t
        return N.Lines(self.n, p(t["wl"]), p(t["gf"]), p(t["cgf"]), p(t["gamma_rad"]),
                       p(t["gamma_stark"]), p(t["gamma_vdw"]), p(t["elo"]), p(t["elem"]),
                       p(t["ion"]), p(t["ltype"]), p(t["process"]), p(t["center_t"]),
                       p(t["center_w"]))


def upload_lines(soa: LineSoA, n_wl: int) -> LineArrays:
    t = {k: to_device(getattr(soa, k)) for k in ("wl", "gf", "cgf", "gamma_rad", "gamma_stark",
                                                  "gamma_vdw", "elo", "elem", "ion", "ltype",
                                                  "process", "center_t", "center_w")}
    corder, cstart = center_csr(soa.center_t, n_wl)
    worder = np.argsort(soa.center_w, kind="stable").astype(np.int64)
    return LineArrays(soa, t, to_device(corder), to_device(cstart), to_device(worder))


class SpectrumPlan:
    """One spectrum's device-resident inputs, buffers and stage launches."""

    def __init__(self, atm: AtmosphereModel, catalog: LineCatalog, wl_start: float, wl_end: float,
                 resolution: float, cutoff: float = 1e-3, microturb_kms: float = 0.0,
                 rhoxj: float = 0.0, grid: Optional[np.ndarray] = None, precision: str = "fp64",
                 bfudge: Optional[np.ndarray] = None, helium_wings: bool = True):
        """helium_wings: the engine's helium pass (SynthesisConfig.enable_helium_wings)."""
        torch = _torch()
        self.ctx = N.context()
        if precision not in ("fp64", "fp32"):
            raise ValueError(f"precision must be 'fp64' or 'fp32', not {precision!r}")
        self.precision = precision
        self.grid = N.wavelength_grid(wl_start, wl_end, resolution) if grid is None else np.asarray(grid)
        self.W = int(self.grid.size)
        self.D = atm.layers
        check_wing_reach(self.grid, atm.doppler_per_ion)
        self.cutoff = float(cutoff)
        self.microturb = float(microturb_kms)
        self.rhoxj = float(rhoxj)
        self.atm = atm
        self.depth = upload_atmosphere(atm)
        self.soa = build_line_soa(catalog, self.grid, self.depth.n_elem,
                                  grid_origin=float(self.grid[0]) if self.W else None)
        self.lines = upload_lines(self.soa, self.W)
        self.hyd = hydrogen_lines(catalog, self.grid)
        self._hyd_args = None
        # fort.19 records (rgfall's routing) of the window, matched to the catalog
        self.f19 = self.he = None
        if atm.population_per_ion is not None:
            pop_shape = np.asarray(atm.population_per_ion).shape
            recs = fort19_records(catalog, wl_start, wl_end, resolution)
            self.f19 = fort19_lines(catalog, recs, self.grid, pop_shape[1], pop_shape[2])
            if helium_wings and atm.doppler_per_ion is not None and pop_shape[2] >= 2:
                self.he = helium_wing_lines(catalog, self.grid, recs, pop_shape[1])
        self._f19_args = None
        self._he_args = None
        f8 = torch.float64
        D, W, L = self.D, self.W, len(self.soa)
        dev = "cuda"
        self.d_grid = to_device(self.grid, np.float64)
        self.cont_abs = torch.empty((D, W), dtype=f8, device=dev)
        self.cont_scat = torch.empty((D, W), dtype=f8, device=dev)
        self.kapmin = torch.empty((D, W), dtype=f8, device=dev)
        self.slinec = torch.empty((D, W), dtype=f8, device=dev)
        self.transp = torch.empty((L, D), dtype=f8, device=dev)
        self.valid = torch.empty((L, D), dtype=torch.uint8, device=dev)
        self.asynth = torch.empty((D, W), dtype=f8, device=dev)
        self.flux_hz = torch.empty(W, dtype=f8, device=dev)
        self.cont_hz = torch.empty(W, dtype=f8, device=dev)
        self.flux = torch.empty(W, dtype=f8, device=dev)
        self.cont = torch.empty(W, dtype=f8, device=dev)
        self.iters = torch.empty(W, dtype=torch.int32, device=dev)
        self.stats = torch.zeros(5, dtype=f8, device=dev)
        rx = np.asarray(atm.depth, np.float64)
        fscat = np.exp(-rx / self.rhoxj) if self.rhoxj > 0.0 else np.zeros_like(rx)
        self.fscat = fscat
        self.one_minus_fscat = to_device((1.0 - fscat)[:, None], np.float64)
        self.fscat_col = to_device(fscat[:, None], np.float64)
        self.line_abs = self.asynth
        self.line_scat = torch.zeros((D, W), dtype=f8, device=dev) if self.rhoxj <= 0.0 else \
            torch.empty((D, W), dtype=f8, device=dev)
        # BFUDGE [D] (physics.bfudge_column; None = 1, LTE) for the line source
        if bfudge is not None:
            bf = np.asarray(bfudge, np.float64).reshape(-1)
            if bf.size != D:
                raise ValueError(f"bfudge has {bf.size} depths, the atmosphere {D}")
            self.bfudge = to_device(bf, np.float64)
        else:
            self.bfudge = None

    # ---------------------------------------------------------------- stages
    def run_continuum(self, stream=None):
        s = stream if stream is not None else N.stream_handle()
        d = self.depth
        N.check("synthe_continuum", self.ctx.lib.synthe_continuum(
            self.ctx.handle, self.D, self.W, N.ptr(self.d_grid), int(d.wledge.shape[0]), N.ptr(d.wledge),
            N.ptr(d.half_edge), N.ptr(d.delta_edge), N.ptr(d.coeff_abs), N.ptr(d.coeff_scat),
            int(d.coeff_log10), N.ptr(self.cont_abs), N.ptr(self.cont_scat), s))
        self.kapmin.copy_(self.cont_abs)
        self.kapmin.add_(self.cont_scat)
        N.check("synthe_source_function", self.ctx.lib.synthe_source_function_bfudge(
            self.ctx.handle, self.D, self.W, N.ptr(self.d_grid), N.ptr(d.temperature), N.ptr(self.bfudge),
            None, None, None, N.ptr(self.slinec), s))

    def run_transp(self, stream=None):
        s = stream if stream is not None else N.stream_handle()
        ls, ds = self.lines.struct(), self.depth.struct()
        N.check("synthe_transp", self.ctx.lib.synthe_transp(
            self.ctx.handle, ct_byref(ls), ct_byref(ds), self.W, N.ptr(self.kapmin), self.cutoff,
            self.microturb, N.ptr(self.transp), N.ptr(self.valid), s))

    def split_scattering(self):
        """ALINE = ASYNTH (1 - FSCAT), SIGMAL = ASYNTH FSCAT (engine/opacity.py:2407-2420)."""
        if self.rhoxj > 0.0:
            self.line_scat.copy_(self.asynth * self.fscat_col)
            self.line_abs = self.asynth * self.one_minus_fscat

    def run_asynth(self, stream=None, stats: bool = False, split: bool = True):
        s = stream if stream is not None else N.stream_handle()
        ls, ds = self.lines.struct(), self.depth.struct()
        self.ctx.set_profile_precision(self.precision)
        N.check("synthe_asynth", self.ctx.lib.synthe_asynth(
            self.ctx.handle, ct_byref(ls), ct_byref(ds), self.W, N.ptr(self.d_grid),
            N.ptr(self.transp), N.ptr(self.valid), N.ptr(self.kapmin), self.cutoff,
            N.ptr(self.lines.corder), N.ptr(self.lines.cstart), N.ptr(self.lines.worder),
            N.ptr(self.asynth), N.ptr(self.stats) if stats else None, s))
        if split:
            self.split_scattering()

    def run_line_opacity(self, stream=None, stats: bool = False):
        """TRANSP and ASYNTH (engine/opacity.py:2320-2363): one fused call
        (TRANSP in the wing set-up's pass) without microturbulence, else the
        two drop-in stages."""
        if self.microturb > 0.0:
            self.run_transp(stream)
            self.run_asynth(stream, stats, split=False)
            self.run_fort19(stream)
            self.run_hydrogen(stream)
            self.run_helium(stream)
            self.split_scattering()
            return
        s = stream if stream is not None else N.stream_handle()
        ls, ds = self.lines.struct(), self.depth.struct()
        self.ctx.set_profile_precision(self.precision)
        N.check("synthe_transp_asynth", self.ctx.lib.synthe_transp_asynth(
            self.ctx.handle, ct_byref(ls), ct_byref(ds), self.W, N.ptr(self.d_grid), N.ptr(self.kapmin),
            self.cutoff, N.ptr(self.lines.corder), N.ptr(self.lines.cstart), N.ptr(self.lines.worder),
            N.ptr(self.transp), N.ptr(self.valid), N.ptr(self.asynth), N.ptr(self.stats) if stats else None, s))
        self.run_fort19(stream)
        self.run_hydrogen(stream)
        self.run_helium(stream)
        self.split_scattering()

    def run_fort19(self, stream=None):
        """The fort.19 records' terms into ASYNTH (engine/opacity.py:2366-2387,
        before the hydrogen lines): normal lines with NBLO/NBUP, autoionising
        lines and merged-continuum ramps, times STIM."""
        f = self.f19
        a = self.atm
        if f is None or len(f) == 0:
            return
        if self._f19_args is None:
            c = np.ascontiguousarray
            pop = c(a.population_per_ion, np.float64)
            emerge, emerge_h = merge_energies(a)
            rho = np.zeros(self.D) if a.mass_density is None else c(a.mass_density, np.float64)
            keep = dict(pop=pop, rho=rho, xne=c(a.electron_density, np.float64), txnxn=c(depth_txnxn(a)),
                        temp=c(a.temperature, np.float64), emerge=c(emerge), emerge_h=c(emerge_h),
                        hckt=c(depth_hckt(a)), vscale=c(depth_doppler_scale(a, self.microturb)))
            fl = N.F19Lines(len(f), *(x.ctypes.data for x in (f.kind, f.wavelength, f.center, f.gf, f.gamma_rad,
                                                             f.gamma_stark, f.gamma_vdw, f.elem, f.ion, f.nelion,
                                                             f.nlast, f.cont_val, f.cat_wavelength, f.elo)))
            fd = N.F19Depths(self.D, pop.shape[1], pop.shape[2], *(keep[k].ctypes.data for k in (
                "pop", "rho", "xne", "txnxn", "temp", "emerge", "emerge_h")), None, keep["hckt"].ctypes.data,
                None, keep["vscale"].ctypes.data)
            self._f19_args = (keep, fl, fd)
        _, fl, fd = self._f19_args
        s = stream if stream is not None else N.stream_handle()
        N.check("synthe_fort19_lines", self.ctx.lib.synthe_fort19_lines(
            self.ctx.handle, ct_byref(fl), ct_byref(fd), self.cutoff, self.W, N.ptr(self.d_grid),
            N.ptr(self.kapmin), None, N.ptr(self.asynth), s))

    def run_helium(self, stream=None):
        """The engine's helium pass (engine/opacity.py:2534-2546, 3551-3590,
        3651-3659): each He line's Voigt wings (the TYPE < -2 branch of
        _apply_fort19_profile) summed per depth, times STIM, into the line
        opacity after the hydrogen lines."""
        h = self.he
        a = self.atm
        if h is None or len(h) == 0:
            return
        if self._he_args is None:
            c = np.ascontiguousarray
            pop = c(a.population_per_ion, np.float64)
            dop = np.asarray(a.doppler_per_ion, np.float64)
            emerge, emerge_h = merge_energies(a)
            rho = np.ones(self.D) if a.mass_density is None else c(a.mass_density, np.float64)
            dopple = c(np.stack([dop[:, h.ion[j] - 1, h.elem[j]] for j in range(len(h))]), np.float64)
            keep = dict(pop=pop, rho=rho, xne=c(a.electron_density, np.float64), txnxn=c(depth_txnxn(a)),
                        temp=c(a.temperature, np.float64), emerge=c(emerge), emerge_h=c(emerge_h),
                        hckt=c(depth_hckt(a)), dopple=dopple)
            fl = N.F19Lines(len(h), *(x.ctypes.data for x in (h.kind, h.wavelength, h.center, h.gf, h.gamma_rad,
                                                             h.gamma_stark, h.gamma_vdw, h.elem, h.ion, h.nelion,
                                                             h.nlast, h.cont_val, h.cat_wavelength, h.elo)))
            fd = N.F19Depths(self.D, pop.shape[1], pop.shape[2], *(keep[k].ctypes.data for k in (
                "pop", "rho", "xne", "txnxn", "temp", "emerge", "emerge_h")), None, keep["hckt"].ctypes.data,
                keep["dopple"].ctypes.data, None)
            self._he_args = (keep, fl, fd)
        _, fl, fd = self._he_args
        s = stream if stream is not None else N.stream_handle()
        N.check("synthe_fort19_lines", self.ctx.lib.synthe_fort19_lines(
            self.ctx.handle, ct_byref(fl), ct_byref(fd), self.cutoff, self.W, N.ptr(self.d_grid),
            N.ptr(self.kapmin), None, N.ptr(self.asynth), s))

    def run_hydrogen(self, stream=None):
        """AHLINE into ASYNTH (engine/opacity.py:2465-2487): the hydrogen lines'
        HPROF4 profiles, skipped as the reference skips them when the bundle
        lacks populations, Doppler widths or mass density."""
        a = self.atm
        if len(self.hyd) == 0 or a.population_per_ion is None or a.doppler_per_ion is None or \
                a.mass_density is None or np.asarray(a.population_per_ion).shape[2] < 1:
            return
        if self._hyd_args is None:
            f = lambda v: None if v is None else np.ascontiguousarray(v, np.float64)  # noqa: E731
            keep = dict(temperature=f(a.temperature), electron_density=f(a.electron_density),
                        mass_density=f(a.mass_density), hckt=f(depth_hckt(a)),
                        pop_h1=f(np.asarray(a.population_per_ion)[:, 0, 0]),
                        dop_h1=f(np.asarray(a.doppler_per_ion)[:, 0, 0]), xnf_he1=f(a.xnf_he1), xnf_h2=f(a.xnf_h2),
                        xnfph=None if getattr(a, "xnfph", None) is None else
                        np.ascontiguousarray(np.asarray(a.xnfph, np.float64).reshape(self.D, -1)[:, :2]),
                        dopph=f(getattr(a, "dopph", None)), turbulent_velocity=f(a.turbulent_velocity))
            h = self.hyd
            hl = N.HLines(len(h), *(x.ctypes.data for x in (h.wavelength, h.center, h.gf, h.elo, h.n_lower,
                                                             h.n_upper)))
            hd = N.HDepths(self.D, *(None if keep[k] is None else keep[k].ctypes.data
                                     for k in ("temperature", "electron_density", "mass_density", "hckt", "pop_h1",
                                               "dop_h1", "xnf_he1", "xnf_h2", "xnfph", "dopph",
                                               "turbulent_velocity")))
            self._hyd_args = (keep, hl, hd)
        _, hl, hd = self._hyd_args
        s = stream if stream is not None else N.stream_handle()
        N.check("synthe_hydrogen_lines", self.ctx.lib.synthe_hydrogen_lines(
            self.ctx.handle, ct_byref(hl), ct_byref(hd), self.microturb, self.cutoff, self.W, N.ptr(self.d_grid),
            N.ptr(self.kapmin), None, N.ptr(self.asynth), s))

    def run_rt(self, stream=None):
        s = stream if stream is not None else N.stream_handle()
        d = self.depth
        N.check("synthe_lte_spectrum", self.ctx.lib.synthe_lte_spectrum(
            self.ctx.handle, self.D, self.W, N.ptr(self.d_grid), N.ptr(d.temperature), N.ptr(d.rhox),
            N.ptr(self.cont_abs), N.ptr(self.cont_scat), N.ptr(self.line_abs), N.ptr(self.line_scat),
            N.ptr(self.slinec), N.ptr(self.flux_hz), N.ptr(self.cont_hz), N.ptr(self.iters), s))
        self.flux.copy_(self.flux_hz)
        self.cont.copy_(self.cont_hz)
        for t in (self.flux, self.cont):
            N.check("synthe_flux_to_lambda", self.ctx.lib.synthe_flux_to_lambda(
                self.ctx.handle, self.W, N.ptr(self.d_grid), N.ptr(t), s))

    def run(self, stats: bool = False):
        if self.W == 0:
            return
        self.run_continuum()
        if len(self.soa) > 0:
            self.run_line_opacity(stats=stats)
        else:
            self.asynth.zero_()
            self.run_fort19()
            self.run_hydrogen()
            self.run_helium()
            self.split_scattering()
        self.run_rt()

    def result(self):
        return dict(wavelength=self.grid.copy(), flux=self.flux.cpu().numpy(),
                    continuum=self.cont.cpu().numpy(), flux_hz=self.flux_hz.cpu().numpy(),
                    cont_hz=self.cont_hz.cpu().numpy())


def ct_byref(s):
    import ctypes

    return ctypes.byref(s)


class BatchPlan:
    """Many spectra -- atmospheres of different depth counts, different
    wavelength windows -- planned once and run as one unit on the GPU.

    ``instances`` holds (atmosphere, catalog, wl_start, wl_end) tuples; each
    catalog is DELLIM-filtered to its window here.  Every instance runs the
    same forward pass as SpectrumPlan (LTE, no line scattering) with the same
    kernels, so results equal per-instance SpectrumPlan results bit for bit;
    the wing gather of all instances is one persistent launch.  ``autotune``
    times the gather's tile and grid candidates on this batch and keeps the
    fastest; ``graph`` captures the whole batch into a hipGraph that run()
    replays.  Inputs stay resident in HBM for the plan's life.
    """

    def __init__(self, instances: Sequence[Tuple[AtmosphereModel, LineCatalog, float, float]],
                 resolution: float = 300000.0, cutoff: float = 1e-3, precision: str = "fp64",
                 autotune: bool = True, graph: bool = True,
                 depth_arrays: Optional[Sequence[DepthArrays]] = None, iters: bool = False):
        """depth_arrays (optional): the device arrays of each instance's
        atmosphere, as upload_atmosphere makes them, for callers that keep
        them elsewhere (e.g. views into one batched tensor they rewrite
        between replays); by default each atmosphere is uploaded here.
        iters: also keep each wavelength's JOSH iteration count (+1000 where
        the total solve took the label-401 branch) in items[i]["iters"]."""
        import ctypes

        from .lines import filter_by_range

        torch = _torch()
        self.ctx = N.context()
        if precision not in ("fp64", "fp32"):
            raise ValueError(f"precision must be 'fp64' or 'fp32', not {precision!r}")
        self.precision = precision
        self.items: List[dict] = []
        arr = (N.Instance * len(instances))()
        for i, (atm, catalog, wl0, wl1) in enumerate(instances):
            grid = N.wavelength_grid(wl0, wl1, resolution)
            check_wing_reach(grid, atm.doppler_per_ion)
            depth = upload_atmosphere(atm) if depth_arrays is None else depth_arrays[i]
            cat = filter_by_range(catalog, wl0, wl1)
            if len(hydrogen_lines(cat, grid)) or len(fort19_records(cat, wl0, wl1, resolution)):
                raise NotImplementedError(f"instance {i}: a batch takes plain lines only; hydrogen and fort.19 "
                                          "lines run in SpectrumPlan")
            soa = build_line_soa(cat, grid, depth.n_elem, grid_origin=float(grid[0]) if grid.size else None)
            lines = upload_lines(soa, grid.size)
            d_grid = to_device(grid, np.float64)
            flux = torch.empty(grid.size, dtype=torch.float64, device="cuda")
            cont = torch.empty(grid.size, dtype=torch.float64, device="cuda")
            its = torch.zeros(grid.size, dtype=torch.int32, device="cuda") if iters else None
            p = N.ptr
            arr[i] = N.Instance(lines.struct(), depth.struct(), int(grid.size), p(d_grid),
                                int(depth.wledge.shape[0]), p(depth.wledge), p(depth.half_edge),
                                p(depth.delta_edge), p(depth.coeff_abs), p(depth.coeff_scat),
                                int(depth.coeff_log10), p(lines.corder), p(lines.cstart), p(lines.worder),
                                float(cutoff), p(flux), p(cont), p(its))
            self.items.append(dict(grid=grid, depth=depth, soa=soa, lines=lines, d_grid=d_grid, flux=flux,
                                   cont=cont, iters=its, wl_start=wl0, wl_end=wl1))
        self._arr = arr
        flags = (N.SYNTHE_BATCH_AUTOTUNE if autotune else 0) | (N.SYNTHE_BATCH_GRAPH if graph else 0)
        self.ctx.set_profile_precision(precision)
        torch.cuda.synchronize()
        h = ctypes.c_void_p()
        N.check("synthe_batch_create", self.ctx.lib.synthe_batch_create(
            self.ctx.handle, len(instances), arr, flags, ctypes.byref(h)))
        self.handle = h

    def __len__(self) -> int:
        return len(self.items)

    def run(self, stream=None) -> None:
        s = stream if stream is not None else N.stream_handle()
        N.check("synthe_batch_run", self.ctx.lib.synthe_batch_run(self.handle, s))

    def info(self) -> dict:
        import ctypes

        kwp, blocks, nt = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        arena = ctypes.c_double()
        tune = np.zeros(16)
        N.check("synthe_batch_info", self.ctx.lib.synthe_batch_info(
            self.handle, ctypes.byref(kwp), ctypes.byref(blocks), tune.ctypes.data, 16, ctypes.byref(nt),
            ctypes.byref(arena)))
        return dict(pixels_per_lane=kwp.value, tile=64 * kwp.value, gather_workgroups=blocks.value,
                    autotune_ms=tune[:nt.value].tolist(), arena_bytes=arena.value)

    def arrays(self, i: int) -> "N.BatchView":
        """Device pointers of instance i's forward arrays after run() (continuum,
        KAPMIN, line source, ASYNTH, zero line scattering, TRANSP, valid), for
        the per-instance adjoint entry points."""
        import ctypes

        v = N.BatchView()
        N.check("synthe_batch_arrays", self.ctx.lib.synthe_batch_arrays(self.handle, int(i), ctypes.byref(v)))
        return v

    def results(self) -> List[dict]:
        return [dict(wavelength=it["grid"].copy(), flux=it["flux"].cpu().numpy(),
                     continuum=it["cont"].cpu().numpy()) for it in self.items]

    def close(self) -> None:
        if getattr(self, "handle", None):
            self.ctx.lib.synthe_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
