"""Differentiable synthesis: flux(T, abundances, log gf, log gamma_vdW) with
hand-written HIP adjoints behind a torch.autograd.Function.

For one model column (gas pressure, column mass, microturbulence and electron
density held fixed), one line list and one wavelength window:

    model = SpectrumModel(columns, catalog, wl0, wl1)
    flux = model(T, dlog_abund, log_gf, log_gvdw)       # F_lambda [W], torch
    flux.backward(weights)                              # gradients on the inputs

Forward: the layer model below (torch, FP64 on the GPU) gives the bundle --
Saha-Boltzmann populations with partition functions, mass density, Doppler
widths, HCKT, TXNXN, continuum coefficients, and optionally the charge
balance n_e -- and the
HIP library runs the same forward pass as SpectrumPlan (continuum, TRANSP,
ASYNTH at the chosen profile precision, LTE JOSH, F_lambda).

Backward (no tape of the forward): synthe_lte_vjp differentiates both JOSH
branches per wavelength in O(depths) (implicit adjoints of the scattering
fixed point and of the label-401 iteration, colour-seeded dual-number
Jacobians of PARCOE/MAP1/DERIV, INTEG reversed), synthe_asynth_vjp recomputes every
profile value and gives per (line, depth) pair dL/dkappa0, dL/da and
dL/d(Doppler width) through the profile shape; this module chains those to
the bundle tensors (then torch autograd takes the layer model to T and the
abundances), and adds the STIM and Planck dependences on T and the continuum
coefficient derivatives.  Branch choices (KAPMIN gates, MAXSTEP, N10DOP,
table indices, label 401) are held at their forward values; FASTEX's table is
differentiated as exp(-x).  Lines must be plain (ltype 0).
"""

from __future__ import annotations

import math
from typing import Dict, Optional

import numpy as np

from . import _native as N
from . import synthetic as S
from .constants import AMU, C_LIGHT_CM, H_PLANCK, K_BOLTZ
from .device import center_csr, to_device
from .lines import LineCatalog, build_line_soa, filter_by_range

LN10 = math.log(10.0)
C_NM = 2.99792458e17


def _torch():
    import torch

    return torch


class LayerModel:
    """torch FP64 restatement of synthetic.build_bundle: the bundle of each
    layer as a function of its temperature and of the 20 abundance offsets
    (dex), with the other .atm columns -- gas pressure, column mass,
    microturbulence and, by default, the electron density XNE -- held fixed,
    as the bundle converter reads them (tools/convert_atm_to_npz.py:343-356).
    Matches build_bundle to rounding.  With electron_density="charge_balance"
    n_e is instead solved from the charge balance (the same bisection, refined
    by one Newton step whose derivative is the implicit derivative)."""

    def __init__(self, columns: S.AtmColumns, device: str = "cuda", electron_density: str = "columns"):
        if electron_density not in ("columns", "charge_balance"):
            raise ValueError("electron_density must be 'columns' or 'charge_balance'")
        self.ne_mode = electron_density
        torch = _torch()
        f8 = torch.float64
        self.device = device
        self.columns = columns
        self.P = torch.tensor(np.asarray(columns.pressure, np.float64), dtype=f8, device=device)
        self.ne_cols = torch.tensor(np.asarray(columns.xne, np.float64), dtype=f8, device=device)
        self.vt = torch.tensor(np.asarray(columns.vturb, np.float64), dtype=f8, device=device)
        self.rhox = np.asarray(columns.rhox, np.float64)
        self.mh = float(columns.mh)
        self.Z = list(S.ABUNDANCE_Z)
        self.logeps0 = torch.tensor([S.ELEMENTS[z]["logeps"] + (self.mh if z > 2 else 0.0) for z in self.Z],
                                    dtype=f8, device=device)
        self.mass = torch.tensor([S.ELEMENTS[z]["mass"] for z in self.Z], dtype=f8, device=device)
        E = S.WLEDGE.size
        half = 0.5 * (S.WLEDGE[:-1] + S.WLEDGE[1:])
        # per (edge interval, node): absorption shape * jump, Rayleigh cross-section
        kfac = np.zeros((E - 1, 3))
        sray = np.zeros((E - 1, 3))
        for e in range(E - 1):
            for j, (lam, side) in enumerate(((S.WLEDGE[e], +1), (half[e], 0), (S.WLEDGE[e + 1], -1))):
                jump = 1.0
                for edge, jv in S._EDGE_JUMP.items():
                    if lam < edge or (lam == edge and side > 0):
                        jump *= jv
                kfac[e, j] = (lam / 500.0) ** 0.8 * jump
                la = lam * 10.0
                sray[e, j] = 5.799e-13 / la ** 4 + 1.422e-6 / la ** 6 + 2.784 / la ** 8
        self.kfac = torch.tensor(kfac, dtype=f8, device=device)
        self.sray = torch.tensor(sray, dtype=f8, device=device)
        # every element at once: stages 0..ns-1 padded to the largest count
        nz, ns = len(self.Z), 1 + max(len(S.ELEMENTS[z]["chi"]) for z in self.Z)
        chi = np.zeros((nz, ns - 1))
        u0 = np.ones((nz, ns))
        up = np.zeros((nz, ns))
        live = np.zeros((nz, ns), bool)
        for iz, z in enumerate(self.Z):
            c = S.ELEMENTS[z]["chi"]
            chi[iz, :len(c)] = c
            live[iz, :len(c) + 1] = True
            for st, (a, b) in enumerate(S.ELEMENTS[z]["u"][:ns]):
                u0[iz, st], up[iz, st] = a, b
        self.chi = torch.tensor(chi, dtype=f8, device=device)[:, :, None]
        self.u0 = torch.tensor(u0, dtype=f8, device=device)[:, :, None]
        self.up = torch.tensor(up, dtype=f8, device=device)[:, :, None]
        self.live = torch.tensor(live, device=device)[:, :, None]
        self.charge = torch.arange(ns, dtype=f8, device=device)[None, :, None]
        # (element, stage) pairs stored in the bundle: stage < N_ION
        zi, si = np.nonzero(live[:, :S.N_ION])
        self.pair_z = torch.tensor(zi, dtype=torch.int64, device=device)
        self.pair_s = torch.tensor(si, dtype=torch.int64, device=device)
        self.pair_el = torch.tensor(np.asarray(self.Z)[zi] - 1, dtype=torch.int64, device=device)
        self.iz_h, self.iz_he = self.Z.index(1), self.Z.index(2)

    def _partitions(self, T):
        """U [element][stage][depth] (u0 (T / 5000)^p; 1 past an element's table)."""
        return self.u0 * (T / 5000.0)[None, None, :] ** self.up

    def _fractions(self, T, ne, U):
        """Stage fractions [element][stage][depth] (0 past an element's last stage):
        synthetic._saha_fractions for all elements at once."""
        torch = _torch()
        ratio = (S.SAHA_CONST * T ** 1.5 / ne)[None, None, :] * (U[:, 1:] / U[:, :-1]) * torch.exp(
            -self.chi / (S.K_EV * T)[None, None, :])
        logr = torch.cat([torch.zeros_like(U[:, :1]), torch.cumsum(torch.log(ratio), dim=1)], dim=1)
        logr = torch.where(self.live, logr, torch.full_like(logr, -math.inf))
        lr = logr - logr.max(dim=1, keepdim=True).values
        w = torch.exp(lr)
        return w / w.sum(dim=1, keepdim=True)

    def _supply(self, T, ne, ab, ntot):
        fr = self._fractions(T, ne, self._partitions(T))
        return ((ab / ab.sum())[:, None] * (self.charge * fr).sum(dim=1)).sum(dim=0) * (ntot - ne)

    def electron_density(self, T, ab):
        torch = _torch()
        ntot = self.P / (K_BOLTZ * T)
        with torch.no_grad():
            Td, abd, nt = T.detach(), ab.detach(), ntot.detach()
            lo = torch.full_like(Td, -30.0)
            hi = torch.log(nt)
            for _ in range(80):
                mid = 0.5 * (lo + hi)
                ne = torch.exp(mid)
                too_big = ne > self._supply(Td, ne, abd, nt)
                hi = torch.where(too_big, mid, hi)
                lo = torch.where(too_big, lo, mid)
            ne0 = torch.exp(0.5 * (lo + hi))
        # implicit derivative of g(ne) = ne - supply(ne) = 0
        with torch.enable_grad():
            x = ne0.clone().requires_grad_(True)
            g = x - self._supply(Td, x, abd, nt)
            gp = torch.autograd.grad(g.sum(), x)[0].detach()
        g_live = ne0 - self._supply(T, ne0, ab, ntot)
        return ne0 - g_live / gp + (g_live / gp).detach()

    def __call__(self, T, dlogA) -> Dict[str, object]:
        torch = _torch()
        ab = 10.0 ** (self.logeps0 + dlogA - 12.0)
        asum = ab.sum()
        ne = self.ne_cols if self.ne_mode == "columns" else self.electron_density(T, ab)
        nnuc = self.P / (K_BOLTZ * T) - ne
        mu = (ab * self.mass).sum() / asum
        rho = nnuc * mu * AMU
        D = T.shape[0]
        U = self._partitions(T)
        n = (ab / asum)[:, None, None] * nnuc[None, None, :] * self._fractions(T, ne, U)  # [el][stage][D]
        nh1, he1 = n[self.iz_h, 0], n[self.iz_he, 0]
        dopz = torch.sqrt(2.0 * K_BOLTZ * T[None, :] / (self.mass[:, None] * AMU)
                          + (self.vt * self.vt)[None, :]) / C_LIGHT_CM  # [el][D]
        pop = torch.zeros((D, S.N_ION, S.N_ELEMENTS), dtype=T.dtype, device=T.device)
        dop = torch.zeros_like(pop)
        pop[:, self.pair_s, self.pair_el] = (n[self.pair_z, self.pair_s] / U[self.pair_z, self.pair_s]).T
        dop[:, self.pair_s, self.pair_el] = dopz[self.pair_z].T
        hckt = H_PLANCK * C_LIGHT_CM / (K_BOLTZ * T)
        txnxn = (nh1 + 0.42 * he1) * (torch.clamp(T, min=1.0) / 10_000.0) ** 0.3
        kr = 0.25 * (self.P / 1.0e5) ** 0.6 * (T / 5800.0) ** 6.0 + 2.0e-4
        kabs = kr[:, None, None] * self.kfac[None]
        sig = (ne[:, None, None] * S.SIGMA_THOMSON + nh1[:, None, None] * self.sray[None]) / rho[:, None, None]
        cabs = torch.log10(torch.clamp(kabs, min=1e-30))
        cscat = torch.log10(torch.clamp(sig, min=1e-30))
        return dict(pop=pop, dop=dop, rho=rho, xne=ne, txnxn=txnxn, hckt=hckt, cabs=cabs, cscat=cscat)


def continuum_lagrange(grid: np.ndarray):
    """Edge interval e [W] and Lagrange weights [W][3] of k_continuum's
    interpolation in the synthetic bundle's edge table: log10 kappa at a pixel
    is sum_j w_j coeff[e][j]."""
    x = np.asarray(grid, np.float64)
    half = 0.5 * (S.WLEDGE[:-1] + S.WLEDGE[1:])
    e = np.clip(np.searchsorted(np.abs(S.WLEDGE), np.abs(x), side="right") - 1, 0, S.WLEDGE.size - 2)
    left, right = np.abs(S.WLEDGE[e]), np.abs(S.WLEDGE[e + 1])
    h, dl = half[e], 0.5 * (S.WLEDGE[1:] - S.WLEDGE[:-1])[e] ** 2
    dl = np.where(dl == 0.0, 1e-20, dl)
    cw = np.stack([(x - h) * (x - right) / dl, (left - x) * (x - right) * 2.0 / dl,
                   (x - left) * (x - h) / dl], axis=1)
    return e, cw


class SpectrumModel:
    """flux(T, dlog_abund, log_gf, log_gvdw) for one column, line list and window."""

    def __init__(self, columns: S.AtmColumns, catalog: LineCatalog, wl_start: float, wl_end: float,
                 resolution: float = 300000.0, cutoff: float = 1e-3, precision: str = "fp32",
                 electron_density: str = "columns", rhoxj: float = 0.0):
        """rhoxj > 0: line scattering as in SpectrumPlan (engine/opacity.py:2407-2420):
        the line opacity splits into absorption asynth (1 - f) and scattering
        asynth f, f = exp(-rhox / rhoxj)."""
        torch = _torch()
        self.ctx = N.context()
        self.layers = LayerModel(columns, electron_density=electron_density)
        self.precision = precision
        self.cutoff = float(cutoff)
        self.grid = N.wavelength_grid(wl_start, wl_end, resolution)
        self.W = int(self.grid.size)
        self.D = int(np.asarray(columns.temperature).size)
        cat = filter_by_range(catalog, wl_start, wl_end)
        self.soa = build_line_soa(cat, self.grid, S.N_ELEMENTS, grid_origin=float(self.grid[0]))
        if np.any(self.soa.ltype != 0):
            raise ValueError("the adjoint supports plain lines (ltype 0) only")
        L = len(self.soa)
        self.L = L
        soa = self.soa
        self.t = {k: to_device(getattr(soa, k)) for k in ("wl", "gf", "gamma_rad", "gamma_stark", "elo", "elem",
                                                          "ion", "ltype", "process", "center_t", "center_w")}
        self.cgf0 = to_device(soa.cgf)
        self.gvdw0 = to_device(soa.gamma_vdw)
        self.log_gf0 = torch.tensor(np.log10(np.maximum(soa.gf, 1e-300)), dtype=torch.float64, device="cuda")
        self.log_gvdw0 = torch.tensor(np.log10(np.maximum(soa.gamma_vdw, 1e-300)), dtype=torch.float64,
                                      device="cuda")
        corder, cstart = center_csr(soa.center_t, self.W)
        self.corder, self.cstart = to_device(corder), to_device(cstart)
        self.worder = to_device(np.argsort(soa.center_w, kind="stable").astype(np.int64))
        self.d_grid = to_device(self.grid, np.float64)
        self.rhox = to_device(self.layers.rhox, np.float64)
        self.rhoxj = float(rhoxj)
        fscat = np.exp(-self.layers.rhox / self.rhoxj) if self.rhoxj > 0.0 else np.zeros_like(self.layers.rhox)
        self.fscat_col = to_device(fscat[:, None], np.float64)
        self.one_minus_fscat = to_device((1.0 - fscat)[:, None], np.float64)
        half = 0.5 * (S.WLEDGE[:-1] + S.WLEDGE[1:])
        self.wledge = to_device(S.WLEDGE.copy())
        self.half = to_device(half)
        self.delta = to_device(0.5 * (S.WLEDGE[1:] - S.WLEDGE[:-1]) ** 2)
        # Lagrange weights of the continuum interpolation (k_continuum) per pixel
        e, cw = continuum_lagrange(self.grid)
        self.cont_w = torch.tensor(cw, dtype=torch.float64, device="cuda")
        # pixels of one edge interval are contiguous (the grid ascends): (interval, first, end)
        cuts = np.flatnonzero(np.diff(e)) + 1
        starts = np.concatenate([[0], cuts]) if self.W else np.zeros(0, np.int64)
        ends = np.concatenate([cuts, [self.W]]) if self.W else np.zeros(0, np.int64)
        self.cont_segs = [(int(e[a]), int(a), int(b)) for a, b in zip(starts, ends)]
        valid = (soa.process != 0) & (soa.elem >= 0)
        ion = np.where(valid, soa.ion - 1, 0)
        el = np.where(valid, soa.elem, 0)
        self.l_ion = torch.tensor(ion, dtype=torch.int64, device="cuda")
        self.l_el = torch.tensor(el, dtype=torch.int64, device="cuda")
        # species (ion, element) of the lines: one-hot [L][n_species] for the per-species sums
        key = ion * S.N_ELEMENTS + el
        self.species, sid = np.unique(key, return_inverse=True)
        # lines ordered by species, so each species' lines are one slice
        order = np.argsort(sid, kind="stable")
        self.species_order = torch.tensor(order, dtype=torch.int64, device="cuda")
        bounds = np.searchsorted(sid[order], np.arange(self.species.size + 1))
        self.species_slices = [(int(a), int(b)) for a, b in zip(bounds[:-1], bounds[1:])]
        self.species_ion = torch.tensor(self.species // S.N_ELEMENTS, dtype=torch.int64, device="cuda")
        self.species_el = torch.tensor(self.species % S.N_ELEMENTS, dtype=torch.int64, device="cuda")

    def base_parameters(self):
        """(T, dlog_abund, log_gf, log_gvdw) of the unperturbed model."""
        torch = _torch()
        T = torch.tensor(np.asarray(self.layers.columns.temperature, np.float64), dtype=torch.float64,
                         device="cuda")
        return T, torch.zeros(20, dtype=torch.float64, device="cuda"), self.log_gf0.clone(), self.log_gvdw0.clone()

    def __call__(self, T, dlog_abund, log_gf, log_gvdw):
        b = self.layers(T, dlog_abund)
        return _FluxFn.apply(b["pop"], b["dop"], b["rho"], b["xne"], b["txnxn"], b["hckt"], T, b["cabs"],
                             b["cscat"], log_gf, log_gvdw, self)

    # ------------------------------------------------------------ internals
    def _line_split(self, asy):
        """(line absorption, line scattering) [D][W] from asynth, as SpectrumPlan."""
        torch = _torch()
        if self.rhoxj <= 0.0:
            return asy, torch.zeros_like(asy)
        return asy * self.one_minus_fscat, asy * self.fscat_col

    def _structs(self, T, rho, xne, txnxn, hckt, pop, dop, cgf, gvdw):
        p = N.ptr
        t = self.t
        lines = N.Lines(self.L, p(t["wl"]), p(t["gf"]), p(cgf), p(t["gamma_rad"]), p(t["gamma_stark"]), p(gvdw),
                        p(t["elo"]), p(t["elem"]), p(t["ion"]), p(t["ltype"]), p(t["process"]), p(t["center_t"]),
                        p(t["center_w"]))
        depths = N.Depths(self.D, p(T), p(self.rhox), p(rho), p(xne), p(txnxn), p(hckt), p(pop), p(dop),
                          S.N_ION, S.N_ELEMENTS)
        return lines, depths


def _c(x):
    return x.detach().contiguous()


class _FluxFn:
    """torch.autograd.Function (built lazily so the module imports without torch)."""

    _fn = None

    @classmethod
    def apply(cls, *args):
        if cls._fn is None:
            cls._fn = _make_function()
        return cls._fn.apply(*args)


def _make_function():
    import ctypes

    import torch

    class FluxFunction(torch.autograd.Function):
        @staticmethod
        def forward(ctx, pop, dop, rho, xne, txnxn, hckt, T, cabs, cscat, log_gf, log_gvdw, model):
            m = model
            lib, h = m.ctx.lib, m.ctx.handle
            s = N.stream_handle()
            pop, dop, rho, xne, txnxn, hckt, T = map(_c, (pop, dop, rho, xne, txnxn, hckt, T))
            cabs, cscat = _c(cabs), _c(cscat)
            cgf = (m.cgf0 * 10.0 ** (log_gf.detach() - m.log_gf0)).contiguous()
            gvdw = (m.gvdw0 * 10.0 ** (log_gvdw.detach() - m.log_gvdw0)).contiguous()
            D, W, L = m.D, m.W, m.L
            f8 = torch.float64
            ca = torch.empty((D, W), dtype=f8, device="cuda")
            cs = torch.empty_like(ca)
            p = N.ptr
            N.check("synthe_continuum", lib.synthe_continuum(h, D, W, p(m.d_grid), S.WLEDGE.size, p(m.wledge),
                                                             p(m.half), p(m.delta), p(cabs), p(cscat), 1, p(ca),
                                                             p(cs), s))
            kapmin = ca + cs
            slinec = torch.empty_like(ca)
            N.check("synthe_source_function", lib.synthe_source_function(h, D, W, p(m.d_grid), p(T), None, None,
                                                                         None, p(slinec), s))
            lines, depths = m._structs(T, rho, xne, txnxn, hckt, pop, dop, cgf, gvdw)
            transp = torch.empty((L, D), dtype=f8, device="cuda")
            valid = torch.empty((L, D), dtype=torch.uint8, device="cuda")
            asy = torch.empty_like(ca)
            if L > 0:
                N.check("synthe_transp", lib.synthe_transp(h, ctypes.byref(lines), ctypes.byref(depths), W,
                                                           p(kapmin), m.cutoff, 0.0, p(transp), p(valid), s))
                m.ctx.set_profile_precision(m.precision)
                N.check("synthe_asynth", lib.synthe_asynth(h, ctypes.byref(lines), ctypes.byref(depths), W,
                                                           p(m.d_grid), p(transp), p(valid), p(kapmin), m.cutoff,
                                                           p(m.corder), p(m.cstart), p(m.worder), p(asy), None, s))
            else:
                asy.zero_()
            la, ls = m._line_split(asy)
            fnu = torch.empty(W, dtype=f8, device="cuda")
            fc = torch.empty_like(fnu)
            N.check("synthe_lte_spectrum", lib.synthe_lte_spectrum(h, D, W, p(m.d_grid), p(T), p(m.rhox), p(ca),
                                                                   p(cs), p(la), p(ls), p(slinec), p(fnu),
                                                                   p(fc), None, s))
            flux = fnu.clone()
            N.check("synthe_flux_to_lambda", lib.synthe_flux_to_lambda(h, W, p(m.d_grid), p(flux), s))
            ctx.model = m
            ctx.save_for_backward(pop, dop, rho, xne, txnxn, hckt, T, ca, cs, slinec, transp, valid, kapmin, asy,
                                  cgf, gvdw)
            return flux

        @staticmethod
        def backward(ctx, g_flux):
            m = ctx.model
            (pop, dop, rho, xne, txnxn, hckt, T, ca, cs, slinec, transp, valid, kapmin, asy, cgf,
             gvdw) = ctx.saved_tensors
            lib, h = m.ctx.lib, m.ctx.handle
            s = N.stream_handle()
            p = N.ptr
            D, W, L = m.D, m.W, m.L
            wl = m.d_grid
            v = (g_flux.detach() * (C_NM / torch.clamp(wl * wl, min=1e-40))).contiguous()
            g_ca = torch.empty_like(ca)
            g_cs = torch.empty_like(ca)
            g_la = torch.empty_like(ca)
            g_b = torch.empty_like(ca)
            g_src = torch.empty_like(ca)
            la, ls = m._line_split(asy)
            N.check("synthe_lte_vjp", lib.synthe_lte_vjp(h, D, W, p(wl), p(T), p(m.rhox), p(ca), p(cs), p(la),
                                                         p(ls), p(slinec), p(v), p(g_ca), p(g_cs), p(g_la),
                                                         p(g_b), p(g_src), s))
            # dL/dasynth: line scattering enters the solve as continuum
            # scattering does (abtot and alpha, not snubar), so dL/d(line_scat)
            # is the cont_scat output
            g_line = g_la if m.rhoxj <= 0.0 else g_la * m.one_minus_fscat + g_cs * m.fscat_col
            # stimulated emission and Planck: direct T dependences
            Tc = torch.clamp(T, min=1.0)[:, None]
            hnu_k = (H_PLANCK / K_BOLTZ) * (C_NM / wl)[None, :]
            xx = hnu_k / Tc
            ex = torch.exp(-xx)
            stim = 1.0 - ex
            dstim_dT = -ex * xx / Tc
            g_T = (g_line * (asy / stim) * dstim_dT).sum(dim=1)
            Tt = T[:, None]
            xb = hnu_k / Tt
            eb = torch.exp(-xb)
            freq = (C_NM / wl)[None, :]
            B = 1.47439e-2 * (freq / 1.0e15) ** 3 * eb / (1.0 - eb)
            dB_dT = B * (xb / Tt) / (1.0 - eb)
            g_T = g_T + ((g_b + g_src) * dB_dT).sum(dim=1)
            # continuum coefficients (k_continuum: 10 ** (Lagrange in lambda))
            g_cabs = torch.zeros((D, S.WLEDGE.size - 1, 3), dtype=torch.float64, device="cuda")
            g_cscat = torch.zeros_like(g_cabs)
            for src_g, val, out in ((g_ca, ca, g_cabs), (g_cs, cs, g_cscat)):
                ok = (val > 0.0) & (val < 3.4e38)
                gl = torch.where(ok, src_g * val * LN10, torch.zeros_like(val))  # dL/d(log10 kappa) per pixel
                # sum over each edge interval's pixels (a contiguous slice)
                for e, w0, w1 in m.cont_segs:
                    out[:, e, :] += (gl[:, w0:w1, None] * m.cont_w[None, w0:w1, :]).sum(dim=1)
            # line opacity
            g_pop = torch.zeros_like(pop)
            g_dop = torch.zeros_like(dop)
            g_rho = torch.zeros_like(rho)
            g_xne = torch.zeros_like(xne)
            g_txnxn = torch.zeros_like(txnxn)
            g_hckt = torch.zeros_like(hckt)
            g_lgf = torch.zeros(L, dtype=torch.float64, device="cuda")
            g_lgw = torch.zeros_like(g_lgf)
            if L > 0:
                s_k0 = torch.empty((L, D), dtype=torch.float64, device="cuda")
                s_a = torch.empty_like(s_k0)
                s_dv = torch.empty_like(s_k0)
                k0 = torch.empty_like(s_k0)
                lines, depths = m._structs(T, rho, xne, txnxn, hckt, pop, dop, cgf, gvdw)
                g_asy = g_line.contiguous()
                N.check("synthe_asynth_vjp", lib.synthe_asynth_vjp(
                    h, ctypes.byref(lines), ctypes.byref(depths), W, p(wl), p(transp), p(valid), p(kapmin),
                    m.cutoff, p(m.corder), p(m.cstart), p(m.worder), p(g_asy), p(s_k0), p(s_a), p(s_dv), p(k0), s))
                ion, el = m.l_ion, m.l_el
                dv = dop[:, ion, el].T  # [L][D]
                pp = pop[:, ion, el].T
                gam = (m.t["gamma_rad"][:, None] + m.t["gamma_stark"][:, None] * xne[None, :]
                       + gvdw[:, None] * txnxn[None, :])
                a = gam / dv
                gk = k0 * s_k0  # dL/d(ln kappa0)
                g_lgf = LN10 * gk.sum(dim=1)
                g_lgw = LN10 * ((gvdw[:, None] * txnxn[None, :] / dv) * s_a).sum(dim=1)
                safe_pp = torch.where(pp > 0, pp, torch.ones_like(pp))
                gpp = torch.where(pp > 0, gk / safe_pp, torch.zeros_like(gk))
                g_rho = -(gk / rho[None, :]).sum(dim=0)
                g_hckt = -(m.t["elo"][:, None] * gk).sum(dim=0)
                gdv = -gk / dv - a / dv * s_a + s_dv
                g_xne = ((m.t["gamma_stark"][:, None] / dv) * s_a).sum(dim=0)
                g_txnxn = ((gvdw[:, None] / dv) * s_a).sum(dim=0)
                # per-species sums over lines (slices of the species order)
                gpp_s = gpp.index_select(0, m.species_order)
                gdv_s = gdv.index_select(0, m.species_order)
                sp_pop = torch.stack([gpp_s[a:b].sum(dim=0) for a, b in m.species_slices])
                sp_dop = torch.stack([gdv_s[a:b].sum(dim=0) for a, b in m.species_slices])
                g_pop[:, m.species_ion, m.species_el] = sp_pop.T
                g_dop[:, m.species_ion, m.species_el] = sp_dop.T
            return (g_pop, g_dop, g_rho, g_xne, g_txnxn, g_hckt, g_T, g_cabs, g_cscat, g_lgf, g_lgw, None)

    return FluxFunction
