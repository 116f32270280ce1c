"""Abundance recovery from many spectra: Gauss-Newton over batched syntheses.

The reference has no inverse solver; this module is the batched-inverse form
of the hot path.  Each target spectrum comes from its own atmosphere with its
own 20 abundance offsets (dex); the unknowns are those offsets.  Every
Gauss-Newton iteration evaluates all spectra at the current offsets and at one
forward-difference step per abundance by replaying one BatchPlan graph (all
spectra per launch), takes a least-squares step per spectrum from the SVD of
its Jacobian (Levenberg-damped when a step fails), and keeps the step when it
lowers that spectrum's misfit.

Abundances enter the forward pass only through the atmosphere bundle
(synthetic.build_bundle): populations ~ A_z / sum(A), the mean molecular
weight (so rho), N(H I) and N(He I) (continuum scattering, TXNXN).  With the
columns' electron density held fixed, as build_bundle does, those are closed
forms in the offsets; AbundanceResponse evaluates them on the GPU for a batch
of offsets and the plan's device arrays are overwritten in place before each
replay, so no bundle is rebuilt on the host.
"""

from __future__ import annotations

import tempfile
import time
from dataclasses import dataclass
from pathlib import Path
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import lines as L
from . import synthetic as S
from .constants import AMU, K_BOLTZ

# elements with lines in the synthetic lists (Fe I/II, Ti I, Cr I, Ca I, Ni I)
LINE_ELEMENTS = (20, 22, 24, 26, 28)


def _torch():
    import torch

    return torch


class AbundanceResponse:
    """The abundance-dependent arrays of synthetic.build_bundle for a batch of
    offsets delta[B][20] (dex, in synthetic.ABUNDANCE_Z order): populations
    [B][D][N_ION][N_ELEMENTS], mass density [B][D], TXNXN [B][D] and the
    log10 continuum scattering coefficients [B][D][E-1][3].  Temperature, the
    electron density, Doppler widths and the absorption coefficients do not
    depend on the abundances."""

    def __init__(self, cols: S.AtmColumns, device: str = "cuda"):
        torch = _torch()
        f8 = torch.float64
        temp = cols.temperature.astype(np.float64)
        ne = cols.xne.astype(np.float64)
        p = cols.pressure.astype(np.float64)
        D = temp.size
        self.D = D
        zs = S.ABUNDANCE_Z
        # populations per nucleus of each element: fraction in stage s / U_s
        unit = np.zeros((D, S.N_ION, len(zs)))
        fr_h1 = fr_he1 = None
        for i, z in enumerate(zs):
            fr = S._saha_fractions(z, temp, ne)
            if z == 1:
                fr_h1 = fr[0]
            if z == 2:
                fr_he1 = fr[0]
            for s in range(min(fr.shape[0], S.N_ION)):
                unit[:, s, i] = fr[s] / S._partition(z, s, temp)
        ab0 = S.abundances(cols.mh)
        nnuc = p / (K_BOLTZ * temp) - ne
        # continuum scattering at the edge nodes: sigma = (ne sT + N(H I) sR(lam)) / rho
        E = S.WLEDGE.size
        half = 0.5 * (S.WLEDGE[:-1] + S.WLEDGE[1:])
        sr = np.zeros((E - 1, 3))
        for e in range(E - 1):
            for j, lam in enumerate((S.WLEDGE[e], half[e], S.WLEDGE[e + 1])):
                lam_a = lam * 10.0
                sr[e, j] = 5.799e-13 / lam_a ** 4 + 1.422e-6 / lam_a ** 6 + 2.784 / lam_a ** 8
        t = lambda a: torch.tensor(np.ascontiguousarray(a), dtype=f8, device=device)  # noqa: E731
        self.unit = t(unit)
        self.nnuc = t(nnuc)
        self.fr_h1, self.fr_he1 = t(fr_h1), t(fr_he1)
        self.ne_st = t(ne * S.SIGMA_THOMSON)
        self.sr = t(sr)
        self.tfac = t((np.maximum(temp, 1.0) / 10_000.0) ** 0.3)
        self.ab0 = t([ab0[z] for z in zs])
        self.mass = t([S.ELEMENTS[z]["mass"] for z in zs])
        self.zidx = torch.tensor([z - 1 for z in zs], dtype=torch.int64, device=device)
        self.device = device

    def arrays(self, delta) -> Dict[str, object]:
        torch = _torch()
        delta = torch.as_tensor(delta, dtype=torch.float64, device=self.device).reshape(-1, self.ab0.numel())
        B = delta.shape[0]
        ab = self.ab0[None, :] * torch.pow(10.0, delta)
        asum = ab.sum(dim=1, keepdim=True)
        frac = ab / asum                                          # [B][20]
        nz = frac[:, None, :] * self.nnuc[None, :, None]          # [B][D][20]
        pop = torch.zeros((B, self.D, S.N_ION, S.N_ELEMENTS), dtype=torch.float64, device=self.device)
        pop[:, :, :, self.zidx] = nz[:, :, None, :] * self.unit[None]
        mu = (ab * self.mass[None, :]).sum(dim=1, keepdim=True) / asum
        rho = self.nnuc[None, :] * mu * AMU                       # [B][D]
        nh1 = nz[:, :, 0] * self.fr_h1[None, :]
        he1 = nz[:, :, 1] * self.fr_he1[None, :]
        sig = (self.ne_st[None, :, None, None] + nh1[:, :, None, None] * self.sr[None, None]) / rho[:, :, None, None]
        cscat = torch.log10(torch.clamp(sig, min=1e-30))
        txnxn = (nh1 + 0.42 * he1) * self.tfac[None, :]
        return dict(pop=pop, rho=rho, txnxn=txnxn, coeff_scat=cscat)


class BatchResponse:
    """AbundanceResponse of many atmospheres at once, written in place.

    The per-atmosphere constants are padded to the deepest atmosphere
    ([atmosphere][depth], padding rows hold harmless values) and
    ``write(delta)`` evaluates the same closed forms as
    AbundanceResponse.arrays, element for element, for every atmosphere in one
    set of batched operations, into caller-owned tensors pop [n][Dmax][ion][Z],
    rho / txnxn [n][Dmax] and coeff_scat [n][Dmax][E-1][3] (rows past an
    atmosphere's depth count are not meaningful)."""

    def __init__(self, responses: Sequence[AbundanceResponse]):
        torch = _torch()
        self.n = len(responses)
        self.dmax = max(r.D for r in responses)
        dev = responses[0].device

        def pad(get, fill=0.0):
            first = get(responses[0])
            out = torch.full((self.n, self.dmax) + tuple(first.shape[1:]), fill, dtype=torch.float64, device=dev)
            for i, r in enumerate(responses):
                out[i, :r.D] = get(r)
            return out

        self.unit = pad(lambda r: r.unit)
        self.nnuc = pad(lambda r: r.nnuc, 1.0)
        self.fr_h1 = pad(lambda r: r.fr_h1)
        self.fr_he1 = pad(lambda r: r.fr_he1)
        self.ne_st = pad(lambda r: r.ne_st)
        self.tfac = pad(lambda r: r.tfac)
        self.ab0 = torch.stack([r.ab0 for r in responses])      # [n][20]
        self.mass = responses[0].mass
        self.sr = responses[0].sr
        self.zidx = responses[0].zidx

    def arrays(self, delta) -> Dict[str, object]:
        """The same closed forms as write(), as new tensors (differentiable in
        delta [n][20]): pop [n][Dmax][ion][Z], rho and txnxn [n][Dmax],
        coeff_scat [n][Dmax][E-1][3]."""
        torch = _torch()
        ab = self.ab0 * torch.pow(10.0, delta)
        asum = ab.sum(dim=1, keepdim=True)
        nz = (ab / asum)[:, None, :] * self.nnuc[:, :, None]
        popz = nz[:, :, None, :] * self.unit                        # [n][D][ion][20]
        mu = (ab * self.mass[None, :]).sum(dim=1, keepdim=True) / asum
        r = self.nnuc * mu * AMU
        nh1 = nz[:, :, 0] * self.fr_h1
        he1 = nz[:, :, 1] * self.fr_he1
        sig = (self.ne_st[:, :, None, None] + nh1[:, :, None, None] * self.sr[None, None]) / r[:, :, None, None]
        return dict(popz=popz, rho=r, txnxn=(nh1 + 0.42 * he1) * self.tfac,
                    coeff_scat=torch.log10(torch.clamp(sig, min=1e-30)))

    def write(self, delta, pop, rho, txnxn, coeff_scat) -> None:
        torch = _torch()
        delta = torch.as_tensor(delta, dtype=torch.float64, device=self.ab0.device).reshape(self.n, -1)
        ab = self.ab0 * torch.pow(10.0, delta)
        asum = ab.sum(dim=1, keepdim=True)
        frac = ab / asum                                            # [n][20]
        nz = frac[:, None, :] * self.nnuc[:, :, None]               # [n][D][20]
        pop[:, :, :, self.zidx] = nz[:, :, None, :] * self.unit
        mu = (ab * self.mass[None, :]).sum(dim=1, keepdim=True) / asum
        r = self.nnuc * mu * AMU                                    # [n][D]
        rho.copy_(r)
        nh1 = nz[:, :, 0] * self.fr_h1
        he1 = nz[:, :, 1] * self.fr_he1
        sig = (self.ne_st[:, :, None, None] + nh1[:, :, None, None] * self.sr[None, None]) / r[:, :, None, None]
        torch.log10(torch.clamp(sig, min=1e-30), out=coeff_scat)
        torch.mul(nh1 + 0.42 * he1, self.tfac, out=txnxn)


@dataclass
class Target:
    cols: S.AtmColumns
    catalog: L.LineCatalog
    wl_start: float
    wl_end: float
    delta_true: np.ndarray  # [20] dex
    flux: np.ndarray        # F_lambda of the FP64 plain path at delta_true
    label401: np.ndarray    # [n_wl] bool: the total solve took the label-401 branch


def window_for(wl_start: float, n_wl: int, resolution: float) -> float:
    """End of a window starting at wl_start whose grid has n_wl points."""
    from . import _native as N

    r = np.log1p(1.0 / resolution)
    w1 = wl_start * np.exp((n_wl - 0.5) * r)
    for _ in range(8):
        m = N.wavelength_grid(wl_start, w1, resolution).size
        if m == n_wl:
            return float(w1)
        w1 *= np.exp((n_wl - m) * r)
    raise RuntimeError("could not size the window")


def make_targets(n: int = 64, n_lines: int = 30_000, n_wl: int = 5_000, seed: int = 5,
                 sigma_dex: float = 0.2, wl_range=(400.0, 700.0), resolution: float = 300000.0,
                 cutoff: float = 1e-3) -> List[Target]:
    """n atmospheres drawn as the batched configuration (Teff U(3500, 40000) K,
    log g U(0, 5), 64-80 layers, [M/H] U(-2, 0.5)), each with a window of n_wl
    points starting in wl_range, its own n_lines-line list, and abundance
    offsets N(0, sigma_dex) for the 19 elements other than hydrogen (the
    reference: a common offset of all 20 leaves every spectrum unchanged);
    the target flux is the FP64 plain path (SpectrumPlan) at those offsets."""
    from .device import SpectrumPlan

    rng = np.random.default_rng(seed)
    out = []
    with tempfile.TemporaryDirectory() as td:
        for i in range(n):
            teff = float(rng.uniform(3500.0, 40000.0))
            logg = float(rng.uniform(0.0, 5.0))
            nl = int(rng.integers(64, 81))
            mh = float(rng.uniform(-2.0, 0.5))
            w0 = float(rng.uniform(wl_range[0], wl_range[1]))
            w1 = window_for(w0, n_wl, resolution)
            delta = rng.normal(0.0, sigma_dex, len(S.ABUNDANCE_Z))
            delta[S.ABUNDANCE_Z.index(1)] = 0.0  # hydrogen is the reference (log eps = 12)
            cols = S.make_atm(teff, logg, nl, mh)
            d = S.synthetic_lines(n_lines, w0, w1, seed=1000 + seed * 7919 + i)
            path = Path(td) / f"lines_{i}.gfall"
            S.write_line_list(path, d)
            cat = L.filter_by_range(L.load_catalog(path), w0, w1)
            atm = S.build_bundle(cols, dict(zip(S.ABUNDANCE_Z, delta.tolist())))
            plan = SpectrumPlan(atm, cat, w0, w1, resolution, cutoff=cutoff)
            plan.run()
            out.append(Target(cols, cat, w0, w1, delta, plan.flux.cpu().numpy(),
                              plan.iters.cpu().numpy() >= 1000))
            del plan
    return out


class InverseProblem:
    """Batched Gauss-Newton recovery of each target's 20 abundance offsets.

    One BatchPlan holds every target's spectrum; evaluate(delta) writes the
    bundle arrays of delta[n][20] into the plan's device arrays and replays
    the graph.  The residual of spectrum i is (F - F_target) / median(F_target)
    over the pixels where neither the target nor the model takes the JOSH
    label-401 branch (surface tau > 20, physics/josh_solver.py:700-727): that
    branch's flux differs from the regular solve's by about the median flux,
    so a saturated pixel crossing the switch jumps by that much for an
    arbitrarily small abundance change, and those jumps -- not the smooth
    response -- made up the whole misfit of the targets where they occur."""

    def __init__(self, targets: Sequence[Target], resolution: float = 300000.0, cutoff: float = 1e-3,
                 autotune: bool = True):
        from .device import BatchPlan

        torch = _torch()
        self.targets = list(targets)
        self.n = len(self.targets)
        self.nz = len(S.ABUNDANCE_Z)
        self.h_index = S.ABUNDANCE_Z.index(1)
        self.line_mask = np.array([z in LINE_ELEMENTS for z in S.ABUNDANCE_Z])
        from .device import upload_atmosphere

        inst = [(S.build_bundle(t.cols), t.catalog, t.wl_start, t.wl_end) for t in self.targets]
        self.resp = [AbundanceResponse(t.cols) for t in self.targets]
        self.batch_resp = BatchResponse(self.resp)
        # the abundance-dependent arrays of every spectrum are views into one
        # padded tensor each, rewritten by one batched evaluation per replay
        depths = [upload_atmosphere(atm) for atm, _, _, _ in inst]
        n, dmax = self.n, self.batch_resp.dmax
        p0, c0 = depths[0].pop, depths[0].coeff_scat
        f8 = torch.float64
        self.pop = torch.zeros((n, dmax) + tuple(p0.shape[1:]), dtype=f8, device="cuda")
        self.rho = torch.ones((n, dmax), dtype=f8, device="cuda")
        self.txnxn = torch.zeros((n, dmax), dtype=f8, device="cuda")
        self.coeff_scat = torch.zeros((n, dmax) + tuple(c0.shape[1:]), dtype=f8, device="cuda")
        for i, dp in enumerate(depths):
            D = dp.n
            for name, big in (("pop", self.pop), ("rho", self.rho), ("txnxn", self.txnxn),
                              ("coeff_scat", self.coeff_scat)):
                view = big[i, :D]
                view.copy_(getattr(dp, name))
                setattr(dp, name, view)
        self.plan = BatchPlan(inst, resolution=resolution, cutoff=cutoff, precision="fp64", autotune=autotune,
                              graph=True, depth_arrays=depths, iters=True)
        self.t_regular = [torch.tensor(~t.label401, device="cuda") for t in self.targets]
        self.valid = None  # per spectrum: pixels regular in target and model at the last evaluation
        self.f_target = [torch.tensor(t.flux, dtype=torch.float64, device="cuda") for t in self.targets]
        # residuals in units of each target's median flux (saturated line
        # cores would dominate relative residuals)
        self.f_scale = [float(np.median(t.flux)) for t in self.targets]
        self.evaluations = 0
        self.cutoff = float(cutoff)
        self.last_delta = None

    def evaluate(self, delta: np.ndarray) -> List[object]:
        """Fluxes (torch, F_lambda) of every spectrum at offsets delta[n][20];
        self.valid gets each spectrum's pixels regular in target and model."""
        torch = _torch()
        self.last_delta = np.array(delta, dtype=np.float64, copy=True)
        self.batch_resp.write(delta, self.pop, self.rho, self.txnxn, self.coeff_scat)
        self.plan.run()
        self.evaluations += 1
        torch.cuda.synchronize()
        self.valid = [(it["iters"] < 1000) & tr for it, tr in zip(self.plan.items, self.t_regular)]
        return [it["flux"].clone() for it in self.plan.items]

    def _vjp_setup(self):
        """Per-spectrum constants of the adjoint chain, built once: the
        continuum interpolation as a [3(E-1)][W] matrix and each spectrum's
        lines ordered by species of the abundance elements (one slice each)."""
        from .adjoint import continuum_lagrange

        torch = _torch()
        f8 = torch.float64
        if getattr(self, "_geo", None) is not None:
            return
        zpos = {z - 1: k for k, z in enumerate(S.ABUNDANCE_Z)}
        E3 = 3 * (S.WLEDGE.size - 1)
        self._geo = []
        for it in self.plan.items:
            ln, grid = it["lines"], it["grid"]
            e, cw = continuum_lagrange(grid)
            M = np.zeros((grid.size, E3))
            for j in range(3):
                M[np.arange(grid.size), 3 * e + j] = cw[:, j]
            soa = ln.soa
            live = (soa.process != 0) & (soa.elem >= 0)
            ion = np.where(live, soa.ion - 1, 0)
            el = np.where(live, soa.elem, 0)
            key = np.where(live, ion * S.N_ELEMENTS + el, -1)
            species = [k for k in np.unique(key) if k >= 0 and (k % S.N_ELEMENTS) in zpos]
            # lines of those species ordered by species: each species one slice
            sid = np.full(key.size, -1)
            for r, k in enumerate(species):
                sid[key == k] = r
            order = np.nonzero(sid >= 0)[0][np.argsort(sid[sid >= 0], kind="stable")]
            bounds = np.searchsorted(sid[order], np.arange(len(species) + 1))
            t = lambda a, dt=f8: torch.tensor(np.ascontiguousarray(a), dtype=dt, device="cuda")  # noqa: E731
            self._geo.append(dict(
                Mt=t(M.T), sp_order=t(order, torch.int64), sp_slices=list(zip(bounds[:-1].tolist(), bounds[1:].tolist())),
                sp_ion=t([k // S.N_ELEMENTS for k in species], torch.int64),
                sp_z=t([zpos[k % S.N_ELEMENTS] for k in species], torch.int64), live=t(live, torch.bool),
                ion=t(ion, torch.int64), el=t(el, torch.int64)))

    def gradient(self, res, masks) -> np.ndarray:
        """dL/d delta [n][20] of L = 1/2 sum_i sum_w (m_i r_i)^2 at the offsets of
        the last evaluate(), by the batched adjoint: for every spectrum the RT
        and line-opacity adjoints (synthe_lte_vjp, synthe_asynth_vjp) run on the
        batch's own forward arrays (synthe_batch_arrays): the RT adjoints of all
        spectra in one batched call (synthe_lte_vjp_batch, one launch per
        stage over every spectrum's wavelengths), then each spectrum's line
        adjoint; their outputs are contracted to
        dL/d(populations, rho, TXNXN) and dL/d(continuum scattering), and one
        autograd pass through the closed-form bundle response
        (BatchResponse.arrays) takes them to the offsets.  Branches (KAPMIN
        gates, MAXSTEP, table indices, label 401) are held at their forward
        values, as in SpectrumModel."""
        import ctypes

        from . import _native as N
        from .adjoint import C_NM

        torch = _torch()
        f8 = torch.float64
        self._vjp_setup()
        lib, h = self.plan.ctx.lib, self.plan.ctx.handle
        st = N.stream_handle()
        p = N.ptr
        dvar = torch.tensor(self.last_delta, dtype=f8, device="cuda", requires_grad=True)
        arr = self.batch_resp.arrays(dvar)
        E3 = 3 * (S.WLEDGE.size - 1)
        # the RT adjoints of every spectrum in one batched call (one launch per stage)
        items = self.plan.items
        views = [self.plan.arrays(i) for i in range(len(items))]
        outs = []
        vi = (N.VjpInstance * len(items))()
        for i, it in enumerate(items):
            dp, wl = it["depth"], it["d_grid"]
            D, W = dp.n, it["grid"].size
            v = views[i]
            g_flam = (res[i] * masks[i].to(f8)) / self.f_scale[i]
            vnu = (g_flam * (C_NM / torch.clamp(wl * wl, min=1e-40))).contiguous()
            g = [torch.empty((D, W), dtype=f8, device="cuda") for _ in range(5)]  # ca, cs, la, planck, source
            outs.append((vnu, g))
            vi[i] = N.VjpInstance(D, W, p(wl), p(dp.temperature), p(dp.rhox), v.cont_abs, v.cont_scat, v.asynth,
                                  v.line_scat, v.slinec, p(vnu), *(p(x) for x in g))
        N.check("synthe_lte_vjp_batch", lib.synthe_lte_vjp_batch(h, len(items), vi, st))
        terms = []
        for i, it in enumerate(items):
            geo = self._geo[i]
            dp, ln = it["depth"], it["lines"]
            D, W, L = dp.n, it["grid"].size, len(ln.soa)
            v = views[i]
            wl = it["d_grid"]
            _, (g_ca, g_cs, g_la, g_b, g_src) = outs[i]
            # continuum scattering: d/d coeff of 10^(sum_j w_j coeff_j), through a surrogate
            lcs = arr["coeff_scat"][i, :D].reshape(D, E3) @ geo["Mt"]
            cs = torch.pow(10.0, torch.clamp(lcs, max=38.0))
            term = (g_cs * cs).sum()
            if L > 0:
                s_k0, s_a, s_dv, k0 = (torch.empty((L, D), dtype=f8, device="cuda") for _ in range(4))
                ls, ds = ln.struct(), dp.struct()
                N.check("synthe_asynth_vjp", lib.synthe_asynth_vjp(
                    h, ctypes.byref(ls), ctypes.byref(ds), W, p(wl), v.transp, v.valid, v.kapmin, self.cutoff,
                    p(ln.corder), p(ln.cstart), p(ln.worder), p(g_la), p(s_k0), p(s_a), p(s_dv), p(k0), st))
                live, ion, el = geo["live"], geo["ion"], geo["el"]
                pp = dp.pop[:, ion, el].T                       # [L][D]
                dv = dp.dop[:, ion, el].T
                gk = torch.where(live[:, None], k0 * s_k0, torch.zeros_like(s_k0))  # dL/d ln KAPPA0
                gpp = torch.where(pp > 0, gk / torch.where(pp > 0, pp, torch.ones_like(pp)), torch.zeros_like(gk))
                # KAPPA0 ~ pop / rho; the damping a ~ gamma_W TXNXN / DOPPLE
                g_rho = -(gk / dp.rho[None, :]).sum(dim=0)
                safe_dv = torch.where(dv > 0, dv, torch.ones_like(dv))
                g_tx = torch.where(live[:, None] & (dv > 0), ln.t["gamma_vdw"][:, None] / safe_dv * s_a,
                                   torch.zeros_like(s_a)).sum(dim=0)
                term = term + (g_rho * arr["rho"][i, :D]).sum() + (g_tx * arr["txnxn"][i, :D]).sum()
                if geo["sp_slices"]:  # per species of the abundance elements
                    gs = gpp.index_select(0, geo["sp_order"])
                    g_sp = torch.stack([gs[a:b].sum(dim=0) for a, b in geo["sp_slices"]])  # [species][D]
                    popz = arr["popz"][i, :D][:, geo["sp_ion"], geo["sp_z"]]              # [D][species]
                    term = term + (g_sp.T * popz).sum()
            terms.append(term)
        g = torch.autograd.grad(torch.stack(terms).sum(), dvar)[0]
        return g.detach().cpu().numpy()

    def residuals(self, flux) -> List[object]:
        """Full residual vectors (F - F_target) / median(F_target); the
        caller masks them with evaluate()'s valid pixels."""
        return [(f - t) / s for f, t, s in zip(flux, self.f_target, self.f_scale)]

    @staticmethod
    def _costs(res, masks=None) -> np.ndarray:
        """Sum of squared residuals per spectrum over masks (one device sync)."""
        torch = _torch()
        if masks is None:
            return torch.stack([(r * r).sum() for r in res]).cpu().numpy()
        return torch.stack([(r * r * m).sum() for r, m in zip(res, masks)]).cpu().numpy()

    def solve(self, max_iter: int = 12, step: float = 1e-3, tol: float = 1e-6, rcond: float = 1e-6,
              cfloor: float = 1e-3, line_floor: float = 2e-3, max_tries: int = 6, max_step: float = 0.5,
              bound: float = 2.0, gradient: str = "differences", jac_every: Optional[int] = None) -> dict:
        """Gauss-Newton from zero offsets, with Levenberg damping on failure.

        Per iteration: a forward-difference Jacobian (one batched evaluation
        per abundance, over the pixels valid at the point and at every step)
        and, per spectrum, the SVD J = U S V^T of its columns; the step is
        -V diag(1 / (s^2 + mu s_0^2)) V^T g with g = J^T r (mu = 0 first:
        plain Gauss-Newton).  gradient="adjoint" takes g from the batched
        adjoint (gradient()) at every point instead, starts from the
        Gauss-Newton Hessian J^T J of the difference Jacobian and updates it
        by BFGS from each accepted step and the change of the adjoint
        gradient, and recomputes the difference Jacobian only after a
        spectrum's trials all failed (or every jac_every iterations when
        given): an iteration then costs the batched forward and adjoint of the
        trial point rather than 20 forwards.
        A spectrum keeps its step if the misfit over the pixels valid at both
        points drops (mu / 10), else
        mu = max(10 mu, 1e-6) and a new trial with the same Jacobian.
        Hydrogen is the reference and stays at zero.  The line elements are
        fitted when their column clears line_floor (in median-flux units per
        dex, the sensitivity report() calls constrained; the difference noise
        is ~1e-4); the other elements, which reach
        the spectrum only through sum(A) and the mean molecular weight, when
        theirs clears cfloor times the largest.  A step is shortened to
        max_step dex, and offsets stay within +-bound."""
        if gradient not in ("differences", "adjoint"):
            raise ValueError("gradient must be 'differences' or 'adjoint'")
        torch = _torch()
        n, nz = self.n, self.nz
        delta = np.zeros((n, nz))
        t0 = time.perf_counter()
        res = self.residuals(self.evaluate(delta))
        valid = list(self.valid)
        cost = self._costs(res, valid)
        history = [float(cost.sum())]
        per_target = [cost.copy()]
        refreshes = 0
        mu = np.zeros(n)
        active = cost > 0
        sens = np.zeros((n, nz))
        marg = np.zeros((n, nz))  # marginal sensitivities 1 / sqrt([(J^T J)^-1]_kk)
        svd = [None] * n
        hess = [None] * n   # adjoint mode: Hessian estimates of the live columns
        last = [None] * n   # adjoint mode: (accepted step, gradient before it)
        common = list(valid)
        refresh = True
        adjoint_calls = 0
        it = 0
        for it in range(1, max_iter + 1):
            if refresh or gradient == "differences" or (jac_every and (it - 1) % jac_every == 0):
                # Jacobian columns over the pixels valid at delta and at every step
                common = [v.clone() for v in valid]
                cols = []
                for k in range(nz):
                    if k == self.h_index:
                        cols.append(None)
                        continue
                    dk = delta.copy()
                    dk[:, k] += step
                    rk = self.residuals(self.evaluate(dk))
                    for i in range(n):
                        common[i] &= self.valid[i]
                    cols.append([(rk[i] - res[i]) / step for i in range(n)])
                refreshes += 1
                svd = []
                for i in range(n):
                    m = common[i].to(torch.float64)
                    idx = [k for k in range(nz) if cols[k] is not None]
                    J = torch.stack([cols[k][i] for k in idx], dim=1) * m[:, None]  # [W][nz - 1]
                    c = torch.linalg.vector_norm(J, dim=0)
                    cn = c.cpu().numpy()
                    live = (cn > cfloor * cn.max()) | ((cn > line_floor) & self.line_mask[idx])
                    sens[i, idx] = cn
                    idx = [k for k, lv in zip(idx, live) if lv]
                    if not idx:
                        svd.append(None)
                        continue
                    u, s_, vh = torch.linalg.svd(J[:, torch.tensor(np.nonzero(live)[0], device=J.device)],
                                                 full_matrices=False)
                    s_np = s_.cpu().numpy()
                    svd.append((idx, s_np, vh.T.cpu().numpy(), s_np * (u.T @ (res[i] * m)).cpu().numpy()))
                    # how far the residual moves per dex of an element when the other live
                    # elements re-adjust: 1 / sigma_k, sigma_k^2 = sum_j V_kj^2 / s_j^2
                    Vn = vh.T.cpu().numpy()
                    with np.errstate(divide="ignore"):
                        var = (Vn ** 2 / np.where(s_np > 0, s_np, 0.0) ** 2).sum(axis=1)
                    marg[i] = 0.0
                    marg[i, idx] = np.where(np.isfinite(var) & (var > 0), 1.0 / np.sqrt(var), 0.0)
                    # Gauss-Newton Hessian J^T J of the live columns, the start of the quasi-Newton updates
                    V_np = vh.T.cpu().numpy()
                    hess[i] = (V_np * s_np ** 2) @ V_np.T
                    last[i] = None
                refresh = False
            if gradient == "adjoint":
                if self.last_delta is None or not np.array_equal(self.last_delta, delta):
                    res = self.residuals(self.evaluate(delta))
                    valid = list(self.valid)
                g = self.gradient(res, valid)
                adjoint_calls += 1
                for i in range(n):
                    if svd[i] is None:
                        continue
                    idx = svd[i][0]
                    gi = g[i, idx]
                    if last[i] is not None:  # BFGS update of the Hessian from the accepted step
                        sk, gk = last[i]
                        yk = gi - gk
                        sy = float(sk @ yk)
                        Hs = hess[i] @ sk
                        sHs = float(sk @ Hs)
                        if sy > 1e-12 * np.linalg.norm(sk) * np.linalg.norm(yk) and sHs > 0:
                            hess[i] = hess[i] - np.outer(Hs, Hs) / sHs + np.outer(yk, yk) / sy
                    lam, Q = np.linalg.eigh(0.5 * (hess[i] + hess[i].T))
                    lam = np.maximum(lam, 0.0)[::-1]
                    Q = Q[:, ::-1]
                    svd[i] = (idx, np.sqrt(lam), Q, Q.T @ gi)
                    last[i] = (np.zeros(len(idx)), gi)
            improved = np.zeros(n, bool)
            for _ in range(max_tries):
                dstep = np.zeros((n, nz))
                for i in np.nonzero(active & ~improved)[0]:
                    if svd[i] is None:
                        continue
                    idx, s_, V, Vg = svd[i]
                    keep = s_ > rcond * s_[0]
                    f = np.zeros_like(s_)
                    f[keep] = 1.0 / (s_[keep] ** 2 + mu[i] * s_[0] ** 2)
                    dstep[i, idx] = -(V @ (f * Vg))
                    big = float(np.abs(dstep[i]).max())
                    if big > max_step:
                        dstep[i] *= max_step / big
                trial = np.clip(delta + dstep, -bound, bound)
                dstep = trial - delta
                rt = self.residuals(self.evaluate(trial))
                vt = self.valid
                base = common if gradient == "differences" else valid
                both = [cm & v for cm, v in zip(base, vt)]  # compare on pixels valid at both points
                ct = self._costs(rt, both)
                c0 = self._costs(res, both)
                cnew = self._costs(rt, vt)
                for i in np.nonzero(active & ~improved)[0]:
                    if ct[i] < c0[i]:
                        improved[i] = True
                        if gradient == "adjoint" and svd[i] is not None and last[i] is not None:
                            last[i] = (dstep[i, svd[i][0]].copy(), last[i][1])
                        delta[i] = trial[i]
                        res[i] = rt[i]
                        valid[i] = vt[i]
                        cost[i] = cnew[i]
                        mu[i] = mu[i] / 10.0 if mu[i] > 1e-9 else 0.0
                        if float(np.abs(dstep[i]).max()) < tol:
                            active[i] = False
                    else:
                        mu[i] = max(10.0 * mu[i], 1e-6)
                if (improved | ~active).all():
                    break
            # a spectrum no trial improves gets a fresh Jacobian at the raised damping, until the
            # damping leaves no step worth taking
            if (active & ~improved).any():
                refresh = True
            active &= improved | (mu < 1e6)
            history.append(float(cost.sum()))
            per_target.append(cost.copy())
            if not active.any():
                break
        elapsed = time.perf_counter() - t0
        return dict(delta=delta, iterations=it, seconds=elapsed, evaluations=self.evaluations,
                    adjoint_evaluations=adjoint_calls, jacobians=refreshes, cost_history=history,
                    cost_per_target=np.array(per_target), cost=cost.copy(), sensitivity=sens, marginal=marg,
                    gradient=gradient)


def report(targets: Sequence[Target], delta: np.ndarray, sensitivity: np.ndarray,
           min_sensitivity: float = 2e-3, marginal: Optional[np.ndarray] = None) -> dict:
    """Abundance errors (dex).  A spectrum constrains a line element (Fe, Ti,
    Cr, Ca, Ni) when its residual moves by at least min_sensitivity (L2, in
    units of the median flux) per dex of that element -- 0.05 dex then changes the misfit
    by 2.5e-3 min_sensitivity^2 or more; neutral-metal lines in hot
    atmospheres are too weak for that.  With `marginal` (solve()'s
    1 / sqrt([(J^T J)^-1]_kk)) the move is the one left when the spectrum's
    other elements re-adjust, so two elements whose lines trade off against
    each other (a fit to 1e-12 of the misfit with both off) are not counted
    as constrained; without it, the Jacobian column's own norm (sensitivity).
    Errors are given for the constrained and the unconstrained pairs, and
    over all 19 offsets (hydrogen is the reference element)."""
    truth = np.stack([t.delta_true for t in targets])
    err = np.abs(delta - truth)
    li = [S.ABUNDANCE_Z.index(z) for z in LINE_ELEMENTS]
    fit = (marginal if marginal is not None else sensitivity)[:, li] >= min_sensitivity
    # the number fraction log(A_z / sum A): what line strengths measure when
    # nothing else in the spectrum pins sum A
    def lsum(d, t):
        ab = S.abundances(t.cols.mh)
        return np.log10(sum(ab[z] * 10.0 ** d[k] for k, z in enumerate(S.ABUNDANCE_Z)))
    ls_est = np.array([lsum(delta[i], t) for i, t in enumerate(targets)])
    ls_true = np.array([lsum(truth[i], t) for i, t in enumerate(targets)])
    err_nf = np.abs((delta[:, li] - ls_est[:, None]) - (truth[:, li] - ls_true[:, None]))
    e_fit, e_not = err[:, li][fit], err[:, li][~fit]
    free = [k for k in range(len(S.ABUNDANCE_Z)) if S.ABUNDANCE_Z[k] != 1]
    out = dict(line_pairs=int(fit.size), line_pairs_constrained=int(fit.sum()),
               constraint="marginal" if marginal is not None else "column",
               line_pairs_column_over_floor=int((sensitivity[:, li] >= min_sensitivity).sum()),
               max_err_constrained_line_dex=float(e_fit.max()) if e_fit.size else None,
               median_err_constrained_line_dex=float(np.median(e_fit)) if e_fit.size else None,
               p90_err_constrained_line_dex=float(np.percentile(e_fit, 90)) if e_fit.size else None,
               rms_err_unconstrained_line_dex=float(np.sqrt((e_not ** 2).mean())) if e_not.size else None,
               max_err_constrained_line_fraction_dex=float(err_nf[fit].max()) if e_fit.size else None,
               median_err_constrained_line_fraction_dex=float(np.median(err_nf[fit])) if e_fit.size else None,
               rms_err_all_19_dex=float(np.sqrt((err[:, free] ** 2).mean())),
               rms_true_offset_dex=float(np.sqrt((truth[:, free] ** 2).mean())))
    return out


def fp64_plain_check(targets: Sequence[Target], delta: np.ndarray, indices: Optional[Sequence[int]] = None,
                     resolution: float = 300000.0, cutoff: float = 1e-3) -> float:
    """Max relative difference between each target and the FP64 plain path
    (SpectrumPlan from a host-built bundle) at the recovered offsets, over the
    pixels with |F| > 1e-3 median(F) (the robust mask of the reference's
    tools/compare_spectra.py:128-135) where neither takes the label-401
    branch (the fit's pixels)."""
    from .device import SpectrumPlan

    worst = 0.0
    for i in (range(len(targets)) if indices is None else indices):
        t = targets[i]
        atm = S.build_bundle(t.cols, dict(zip(S.ABUNDANCE_Z, np.asarray(delta[i]).tolist())))
        plan = SpectrumPlan(atm, t.catalog, t.wl_start, t.wl_end, resolution, cutoff=cutoff)
        plan.run()
        f = plan.flux.cpu().numpy()
        m = (np.abs(t.flux) > 1e-3 * np.median(np.abs(t.flux))) & ~t.label401 & (plan.iters.cpu().numpy() < 1000)
        worst = max(worst, float(np.max(np.abs(f[m] - t.flux[m]) / np.abs(t.flux[m]))))
    return worst
