"""Seeded synthetic workloads: atmospheres (.atm and bundle) and GFALL line lists.

Atmosphere recipe: Rosseland tau log-spaced 1e-7..1e2, grey Eddington
T^4 = 3/4 Teff^4 (tau + 2/3), hydrostatic dP/dtau = g / kappa_R(P, T) with a
power-law mean opacity, LTE Saha ionisation of 20 elements for the electron
density and the line species' populations, solar abundances scaled by [M/H].
The physics is the build's own (the reference computes these offline in
tools/pops_exact.py and physics/kapp.py); only the bundle schema follows the
reference (tools/convert_atm_to_npz.py:2755-2830, io/atmosphere.py:82-506).

Line recipe: lambda uniform in ln lambda; species Fe I 40 %, Fe II 15 %,
Ti I 10 %, Cr I 10 %, Ca I 10 %, Ni I 15 %; log gf ~ N(-2, 1.5) clipped to
[-8, 1]; E_low ~ U(0, 8) eV; log gamma_rad ~ U(7, 9), log gamma_Stark ~
U(-7, -4), log gamma_vdW ~ U(-8, -7); written as GFALL with EP = E + 1e7/lambda.
"""

from __future__ import annotations

from dataclasses import dataclass
from pathlib import Path
from typing import Dict, Optional, Sequence, Tuple

import numpy as np

from .atmosphere import AtmosphereModel
from .constants import AMU, C_LIGHT_CM, H_PLANCK, K_BOLTZ
from . import lines as _lines

EV_CM = 8065.54429  # cm^-1 per eV (atomic.py:15)
K_EV = 8.617333262e-5  # eV / K
SAHA_CONST = 4.829812e15  # 2 (2 pi m_e k / h^2)^1.5, cm^-3 K^-1.5
SIGMA_THOMSON = 6.6524587e-25
N_ION = 6
N_ELEMENTS = 99

# 20 elements: Z, mass (amu), solar log eps, ionisation potentials (eV), partition
# functions U_i(T) = u0 * (T/5000)**p for stages I..III.
ELEMENTS: Dict[int, dict] = {
    1: dict(mass=1.008, logeps=12.00, chi=(13.598,), u=((2.0, 0.0), (1.0, 0.0))),
    2: dict(mass=4.003, logeps=10.93, chi=(24.587, 54.418), u=((1.0, 0.0), (2.0, 0.0), (1.0, 0.0))),
    6: dict(mass=12.011, logeps=8.43, chi=(11.260, 24.383), u=((9.3, 0.05), (6.0, 0.0), (1.0, 0.0))),
    7: dict(mass=14.007, logeps=7.83, chi=(14.534, 29.601), u=((4.0, 0.05), (9.0, 0.02), (6.0, 0.0))),
    8: dict(mass=15.999, logeps=8.69, chi=(13.618, 35.121), u=((8.7, 0.03), (4.0, 0.0), (9.0, 0.0))),
    11: dict(mass=22.990, logeps=6.24, chi=(5.139, 47.286), u=((2.0, 0.1), (1.0, 0.0), (6.0, 0.0))),
    12: dict(mass=24.305, logeps=7.60, chi=(7.646, 15.035), u=((1.0, 0.05), (2.0, 0.0), (1.0, 0.0))),
    13: dict(mass=26.982, logeps=6.45, chi=(5.986, 18.829), u=((5.8, 0.02), (1.0, 0.0), (2.0, 0.0))),
    14: dict(mass=28.086, logeps=7.51, chi=(8.152, 16.346), u=((9.4, 0.05), (5.6, 0.02), (1.0, 0.0))),
    16: dict(mass=32.06, logeps=7.12, chi=(10.360, 23.338), u=((8.5, 0.04), (4.0, 0.0), (9.0, 0.0))),
    19: dict(mass=39.098, logeps=5.03, chi=(4.341, 31.625), u=((2.1, 0.3), (1.0, 0.0), (6.0, 0.0))),
    20: dict(mass=40.078, logeps=6.34, chi=(6.113, 11.872, 50.913), u=((1.2, 0.4), (2.2, 0.1), (1.0, 0.0))),
    21: dict(mass=44.956, logeps=3.15, chi=(6.561, 12.800), u=((10.0, 0.3), (21.0, 0.2), (10.0, 0.0))),
    22: dict(mass=47.867, logeps=4.95, chi=(6.828, 13.576, 27.49), u=((30.0, 0.4), (50.0, 0.3), (20.0, 0.1))),
    23: dict(mass=50.942, logeps=3.93, chi=(6.746, 14.66), u=((40.0, 0.4), (30.0, 0.3), (25.0, 0.1))),
    24: dict(mass=51.996, logeps=5.64, chi=(6.767, 16.486, 30.96), u=((10.0, 0.3), (7.0, 0.1), (20.0, 0.1))),
    25: dict(mass=54.938, logeps=5.43, chi=(7.434, 15.640), u=((6.5, 0.2), (7.5, 0.1), (6.0, 0.0))),
    26: dict(mass=55.845, logeps=7.50, chi=(7.902, 16.199, 30.65), u=((25.0, 0.4), (40.0, 0.3), (25.0, 0.1))),
    27: dict(mass=58.933, logeps=4.99, chi=(7.881, 17.084), u=((30.0, 0.3), (25.0, 0.2), (25.0, 0.0))),
    28: dict(mass=58.693, logeps=6.22, chi=(7.640, 18.169, 35.19), u=((30.0, 0.3), (10.0, 0.1), (20.0, 0.1))),
}
ABUNDANCE_Z = tuple(sorted(ELEMENTS))  # the 20 elements, ascending Z

# continuum edges (nm) and the absorption jump below each edge
WLEDGE = np.array([50.0, 91.18, 150.0, 206.8, 250.0, 364.7, 500.0, 820.6, 1200.0, 1458.0,
                   1644.4, 2000.0, 5000.0], dtype=np.float64)
_EDGE_JUMP = {364.7: 1.6, 820.6: 1.25, 1644.4: 1.3, 206.8: 2.0, 91.18: 5.0}

LINE_SPECIES = ((26.00, 0.40), (26.01, 0.15), (22.00, 0.10), (24.00, 0.10), (20.00, 0.10),
                (28.00, 0.15))


def tau_grid(n_layers: int = 72) -> np.ndarray:
    return np.logspace(-7.0, 2.0, n_layers)


def _partition(z: int, stage: int, temp: np.ndarray) -> np.ndarray:
    us = ELEMENTS[z]["u"]
    u0, p = us[stage] if stage < len(us) else (1.0, 0.0)
    return u0 * (temp / 5000.0) ** p


def _kappa_rosseland(p: np.ndarray, temp: np.ndarray) -> np.ndarray:
    """Power-law Rosseland mean opacity, cm^2 g^-1 (H- like: P^0.6 T^6 scaling)."""
    return 0.25 * (p / 1.0e5) ** 0.6 * (temp / 5800.0) ** 6.0 + 2.0e-4


def _hydrostatic(tau: np.ndarray, temp: np.ndarray, g: float) -> np.ndarray:
    """Integrate dP/dln(tau) = tau g / kappa_R(P, T) with second-order steps."""
    p = np.empty_like(tau)
    pp = g * tau[0] / 2.0e-4
    for _ in range(200):
        pp = g * tau[0] / _kappa_rosseland(np.array([pp]), temp[:1])[0]
    p[0] = pp
    lt = np.log(tau)
    for i in range(1, tau.size):
        h = lt[i] - lt[i - 1]
        f0 = tau[i - 1] * g / _kappa_rosseland(np.array([p[i - 1]]), temp[i - 1:i])[0]
        pe = p[i - 1] + h * f0
        f1 = tau[i] * g / _kappa_rosseland(np.array([pe]), temp[i:i + 1])[0]
        p[i] = p[i - 1] + 0.5 * h * (f0 + f1)
    return p


def abundances(mh: float = 0.0, delta: Optional[Dict[int, float]] = None) -> Dict[int, float]:
    """Number abundance relative to H: 10**(log eps - 12), metals scaled by [M/H]."""
    out = {}
    for z, e in ELEMENTS.items():
        le = e["logeps"] + (mh if z > 2 else 0.0) + (delta or {}).get(z, 0.0)
        out[z] = 10.0 ** (le - 12.0)
    return out


def _saha_fractions(z: int, temp: np.ndarray, ne: np.ndarray) -> np.ndarray:
    """Stage fractions [n_stage][D] of element z (stages I..len(chi)+1)."""
    chi = ELEMENTS[z]["chi"]
    ns = len(chi) + 1
    logr = np.zeros((ns, temp.size))
    for i, x in enumerate(chi):
        u_lo = _partition(z, i, temp)
        u_hi = _partition(z, i + 1, temp)
        ratio = SAHA_CONST * temp ** 1.5 / ne * (u_hi / u_lo) * np.exp(-x / (K_EV * temp))
        logr[i + 1] = logr[i] + np.log(ratio)
    logr -= logr.max(axis=0)
    w = np.exp(logr)
    return w / w.sum(axis=0)


def _electron_density(temp: np.ndarray, p: np.ndarray, ab: Dict[int, float]) -> np.ndarray:
    """Solve charge balance n_e = sum_z sum_i i N_z,i with P = (n_nuc + n_e) k T."""
    ntot = p / (K_BOLTZ * temp)
    asum = sum(ab.values())
    lo = np.full_like(temp, -30.0)
    hi = np.log(ntot)
    for _ in range(80):
        mid = 0.5 * (lo + hi)
        ne = np.exp(mid)
        nnuc = ntot - ne
        supply = np.zeros_like(temp)
        for z in ELEMENTS:
            fr = _saha_fractions(z, temp, ne)
            charge = np.arange(fr.shape[0])[:, None]
            supply += ab[z] / asum * nnuc * (charge * fr).sum(axis=0)
        too_big = ne > supply
        hi = np.where(too_big, mid, hi)
        lo = np.where(too_big, lo, mid)
    return np.exp(0.5 * (lo + hi))


def _continuum_node_values(lam: float, kr: np.ndarray, ne: np.ndarray, nh1: np.ndarray,
                           rho: np.ndarray, side: int) -> tuple:
    """kappa_abs and sigma (cm^2/g) at wavelength lam on one side of an edge."""
    shape = (lam / 500.0) ** 0.8
    jump = 1.0
    for edge, j in _EDGE_JUMP.items():
        if lam < edge or (lam == edge and side > 0):
            jump *= j
    kabs = kr * shape * jump
    lam_a = lam * 10.0
    sig_r = 5.799e-13 / lam_a ** 4 + 1.422e-6 / lam_a ** 6 + 2.784 / lam_a ** 8
    sig = (ne * SIGMA_THOMSON + nh1 * sig_r) / rho
    return kabs, sig


@dataclass
class AtmColumns:
    """The .atm deck columns (tools/convert_atm_to_npz.py:343-356)."""

    rhox: np.ndarray
    temperature: np.ndarray
    pressure: np.ndarray
    xne: np.ndarray
    abross: np.ndarray
    accrad: np.ndarray
    vturb: np.ndarray  # cm/s
    teff: float
    logg: float
    mh: float


def make_atm(teff: float, logg: float, n_layers: int = 72, mh: float = 0.0,
             vturb_kms: float = 1.5) -> AtmColumns:
    """Grey Eddington, hydrostatic, Saha-ionised model atmosphere."""
    tau = tau_grid(n_layers)
    temp = (0.75 * teff ** 4 * (tau + 2.0 / 3.0)) ** 0.25
    g = 10.0 ** logg
    p = _hydrostatic(tau, temp, g)
    ab = abundances(mh)
    ne = _electron_density(temp, p, ab)
    return AtmColumns(rhox=p / g, temperature=temp, pressure=p, xne=ne,
                      abross=_kappa_rosseland(p, temp), accrad=np.zeros(n_layers),
                      vturb=np.full(n_layers, vturb_kms * 1.0e5), teff=teff, logg=logg, mh=mh)


def write_atm(path: Path, atm: AtmColumns) -> None:
    """Kurucz-style .atm deck: header, READ DECK6, RHOX T P XNE ABROSS ACCRAD VTURB."""
    lines = [f"TEFF {atm.teff:8.1f}  GRAVITY {atm.logg:8.5f} LTE",
             "TITLE SYNTHETIC GREY EDDINGTON MODEL",
             f"ABUNDANCE SCALE {10.0 ** atm.mh:9.5f}",
             f"READ DECK6 {atm.rhox.size} RHOX,T,P,XNE,ABROSS,ACCRAD,VTURB"]
    for i in range(atm.rhox.size):
        lines.append(f" {atm.rhox[i]:.8E} {atm.temperature[i]:9.1f} {atm.pressure[i]:.3E}"
                     f" {atm.xne[i]:.3E} {atm.abross[i]:.3E} {atm.accrad[i]:.3E}"
                     f" {atm.vturb[i]:.3E}")
    lines.append("PRADK 0.0")
    Path(path).write_text("\n".join(lines) + "\n")


def read_atm(path: Path) -> AtmColumns:
    """Parse the deck written by write_atm (the DECK6 block of a Kurucz .atm)."""
    text = Path(path).read_text().splitlines()
    teff = logg = 0.0
    mh = 0.0
    rows = []
    it = iter(text)
    for ln in it:
        tok = ln.split()
        if not tok:
            continue
        if tok[0] == "TEFF":
            teff, logg = float(tok[1]), float(tok[3])
        elif tok[0] == "ABUNDANCE" and tok[1] == "SCALE":
            mh = float(np.log10(float(tok[2])))
        elif tok[0] == "READ" and tok[1] == "DECK6":
            n = int(tok[2])
            for _ in range(n):
                rows.append([float(x) for x in next(it).split()[:7]])
    a = np.array(rows, dtype=np.float64)
    return AtmColumns(rhox=a[:, 0], temperature=a[:, 1], pressure=a[:, 2], xne=a[:, 3],
                      abross=a[:, 4], accrad=a[:, 5], vturb=a[:, 6], teff=teff, logg=logg, mh=mh)


def build_bundle(atm: AtmColumns, abund_delta: Optional[Dict[int, float]] = None) -> AtmosphereModel:
    """Derive the hot-path bundle from .atm columns (the build's own bundle builder)."""
    temp = atm.temperature.astype(np.float64)
    ne = atm.xne.astype(np.float64)
    p = atm.pressure.astype(np.float64)
    ab = abundances(atm.mh, abund_delta)
    asum = sum(ab.values())
    nnuc = p / (K_BOLTZ * temp) - ne
    mu = sum(ab[z] * ELEMENTS[z]["mass"] for z in ELEMENTS) / asum
    rho = nnuc * mu * AMU
    D = temp.size
    pop = np.zeros((D, N_ION, N_ELEMENTS), dtype=np.float64)
    dop = np.zeros((D, N_ION, N_ELEMENTS), dtype=np.float64)
    vt = atm.vturb.astype(np.float64)
    xnf = {}
    for z, e in ELEMENTS.items():
        fr = _saha_fractions(z, temp, ne)
        nz = ab[z] / asum * nnuc
        for s in range(fr.shape[0]):
            n_s = nz * fr[s]
            xnf[(z, s)] = n_s
            if s < N_ION:
                pop[:, s, z - 1] = n_s / _partition(z, s, temp)
                dop[:, s, z - 1] = np.sqrt(2.0 * K_BOLTZ * temp / (e["mass"] * AMU) + vt * vt) / C_LIGHT_CM
    kr = atm.abross.astype(np.float64)
    nh1 = xnf[(1, 0)]
    E = WLEDGE.size
    cabs = np.zeros((D, E - 1, 3))
    cscat = np.zeros((D, E - 1, 3))
    half = 0.5 * (WLEDGE[:-1] + WLEDGE[1:])
    for e in range(E - 1):
        for j, (lam, side) in enumerate(((WLEDGE[e], +1), (half[e], 0), (WLEDGE[e + 1], -1))):
            ka, sg = _continuum_node_values(lam, kr, ne, nh1, rho, side)
            cabs[:, e, j] = np.log10(np.maximum(ka, 1e-30))
            cscat[:, e, j] = np.log10(np.maximum(sg, 1e-30))
    return AtmosphereModel(
        depth=atm.rhox.astype(np.float64),
        temperature=temp,
        gas_pressure=p,
        electron_density=ne,
        mass_density=rho,
        turbulent_velocity=vt,
        hckt=H_PLANCK * C_LIGHT_CM / (K_BOLTZ * temp),
        xnf_h=nh1,
        xnf_he1=xnf[(2, 0)],
        xnf_h2=np.zeros(D),
        # H I ground level (g = 2 of the partition function) and protons
        xnfph=np.ascontiguousarray(np.stack([2.0 * pop[:, 0, 0], xnf[(1, 1)]], axis=1)),
        population_per_ion=pop,
        doppler_per_ion=dop,
        continuum_wledge=WLEDGE.copy(),
        continuum_half_edge=half,
        continuum_delta_edge=0.5 * (WLEDGE[1:] - WLEDGE[:-1]) ** 2,
        continuum_abs_coeff=cabs,
        continuum_scat_coeff=cscat,
        continuum_coeff_log10=True,
        metadata={"teff": float(atm.teff), "logg": float(atm.logg), "mh": float(atm.mh)},
    )


def synthetic_lines(n: int, wl0: float, wl1: float, seed: int) -> Dict[str, np.ndarray]:
    """Draw a line list (sorted by wavelength, as GFALL files are)."""
    rng = np.random.default_rng(seed)
    lnw = rng.uniform(np.log(wl0), np.log(wl1), n)
    wl = np.sort(np.exp(lnw))
    codes = np.array([c for c, _ in LINE_SPECIES])
    probs = np.array([p for _, p in LINE_SPECIES])
    code = codes[rng.choice(codes.size, size=n, p=probs / probs.sum())]
    loggf = np.clip(rng.normal(-2.0, 1.5, n), -8.0, 1.0)
    e_low = rng.uniform(0.0, 8.0, n) * EV_CM
    gr = rng.uniform(7.0, 9.0, n)
    gs = rng.uniform(-7.0, -4.0, n)
    gw = rng.uniform(-8.0, -7.0, n)
    return dict(wl=wl, code=code, loggf=loggf, e_low=e_low, gr=gr, gs=gs, gw=gw)


def hydrogen_series(wl0: float, wl1: float, margin_nm: float = 100.0, n_max: int = 6,
                    m_max: int = 40) -> Dict[str, np.ndarray]:
    """H I lines n -> m (n <= n_max, m <= m_max) within margin_nm of the window,
    as GFALL records (code 1.00, NBLO/NBUP set): level energies
    R_H (1 - 1/n^2) cm^-1, vacuum wavelengths from them, Kramers oscillator
    strengths f = 32 / (3 pi sqrt 3) / (n^5 m^3 (1/n^2 - 1/m^2)^3) with a
    Gaunt factor 0.87 and g = 2 n^2."""
    ryd = 109677.576
    rows = []
    for n in range(1, n_max + 1):
        for m in range(n + 1, m_max + 1):
            e_n, e_m = ryd * (1.0 - 1.0 / n ** 2), ryd * (1.0 - 1.0 / m ** 2)
            wl = 1.0e7 / (e_m - e_n)
            if not (wl0 - margin_nm <= wl <= wl1 + margin_nm):
                continue
            f = 0.87 * 32.0 / (3.0 * np.pi * np.sqrt(3.0)) / (n ** 5 * m ** 3 * (1.0 / n ** 2 - 1.0 / m ** 2) ** 3)
            rows.append((wl, np.log10(2.0 * n * n * f), e_n, n, m))
    rows.sort()
    a = np.array(rows, dtype=np.float64).reshape(-1, 5)
    k = a.shape[0]
    return dict(wl=a[:, 0], code=np.full(k, 1.00), loggf=a[:, 1], e_low=a[:, 2], gr=np.full(k, 8.5),
                gs=np.full(k, -5.0), gw=np.full(k, -7.5), n_lower=a[:, 3].astype(np.int64),
                n_upper=a[:, 4].astype(np.int64))


def merge_line_lists(*lists: Dict[str, np.ndarray]) -> Dict[str, np.ndarray]:
    """Concatenate line draws and sort them by wavelength (GFALL order)."""
    keys = ("wl", "code", "loggf", "e_low", "gr", "gs", "gw")
    out = {k: np.concatenate([d[k] for d in lists]) for k in keys}
    out["n_lower"] = np.concatenate([d.get("n_lower", np.zeros(d["wl"].size, np.int64)) for d in lists])
    out["n_upper"] = np.concatenate([d.get("n_upper", np.zeros(d["wl"].size, np.int64)) for d in lists])
    order = np.argsort(out["wl"], kind="stable")
    return {k: v[order] for k, v in out.items()}


def write_line_list(path: Path, d: Dict[str, np.ndarray]) -> None:
    _lines.write_gfall(path, d["wl"], d["loggf"], d["code"], d["e_low"], d["gr"], d["gs"], d["gw"],
                       d.get("n_lower"), d.get("n_upper"))


@dataclass
class Workload:
    atmosphere: AtmosphereModel
    catalog: "_lines.LineCatalog"
    wl_start: float
    wl_end: float
    resolution: float
    name: str
    line_spec: Optional[Tuple[int, float, float, int]] = None  # (n, lambda0, lambda1, seed) drawn


def write_workload_gfall(wk: Workload, path: Path, window: Optional[Tuple[float, float]] = None,
                         pad_nm: float = 2.0) -> Path:
    """Write the GFALL text a workload's catalog was parsed from (the draw is
    seeded, so it is redrawn).  With window, only lines drawn within pad_nm of
    it are written: the DELLIM margin of these species' lines is 0.1 nm times
    max(1, lambda0/500) (< 0.4 nm below 2000 nm) and a parsed wavelength is
    within 1e-3 nm of the drawn one, so the window's filtered catalog is
    unchanged."""
    if wk.line_spec is None:
        raise ValueError("workload has no line draw to write")
    n, lw0, lw1, seed = wk.line_spec
    d = synthetic_lines(n, lw0, lw1, seed)
    if window is not None:
        keep = (d["wl"] >= window[0] - pad_nm) & (d["wl"] <= window[1] + pad_nm)
        d = {k: v[keep] for k, v in d.items()}
    write_line_list(Path(path), d)
    return Path(path)


def make_workload(name: str, teff: float, logg: float, n_lines: int, wl0: float, wl1: float,
                  resolution: float = 300000.0, n_layers: int = 72, seed: int = 1,
                  cache_dir: Optional[Path] = None, line_wl: Optional[Sequence[float]] = None,
                  mh: float = 0.0) -> Workload:
    """Build (or load from cache_dir) an atmosphere bundle and a parsed line catalog.

    The line list is drawn over line_wl (defaults to the synthesis window),
    written as GFALL, parsed back and DELLIM-filtered to the window, exactly as
    the reference's compiler path does with a GFALL file.
    """
    lw0, lw1 = (wl0, wl1) if line_wl is None else line_wl
    cache = Path(cache_dir) if cache_dir else None
    tag = f"{name}_t{teff:g}_g{logg:g}_m{mh:g}_d{n_layers}_n{n_lines}_{lw0:g}-{lw1:g}_s{seed}"
    atm_npz = cat_npz = None
    if cache is not None:
        cache.mkdir(parents=True, exist_ok=True)
        atm_npz, cat_npz = cache / f"{tag}_atm.npz", cache / f"{tag}_lines.npz"
        if atm_npz.exists() and cat_npz.exists():
            from .atmosphere import load_cached
            cat = _lines.LineCatalog.load_npz(cat_npz)
            return Workload(load_cached(atm_npz), _lines.filter_by_range(cat, wl0, wl1),
                            wl0, wl1, resolution, name, (n_lines, lw0, lw1, seed))
    atm = build_bundle(make_atm(teff, logg, n_layers, mh))
    d = synthetic_lines(n_lines, lw0, lw1, seed)
    import tempfile
    with tempfile.TemporaryDirectory() as td:
        gf = Path(td) / "lines.gfall"
        write_line_list(gf, d)
        cat = _lines.load_catalog(gf)
    if cache is not None:
        atm.save(atm_npz)
        cat.save_npz(cat_npz)
    return Workload(atm, _lines.filter_by_range(cat, wl0, wl1), wl0, wl1, resolution, name,
                    (n_lines, lw0, lw1, seed))


def batch_draws(n: int = 256, seed: int = 3):
    """The batched configuration's draws: per instance (Teff ~ U(3500, 40000) K,
    log g ~ U(0, 5), layers ~ U{64..80}, [M/H] ~ U(-2, 0.5), window start ~
    U(300, 1700) nm, width ~ U(5, 50) nm)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        teff = float(rng.uniform(3500.0, 40000.0))
        logg = float(rng.uniform(0.0, 5.0))
        nl = int(rng.integers(64, 81))
        mh = float(rng.uniform(-2.0, 0.5))
        w0 = float(rng.uniform(300.0, 1700.0))
        w1 = w0 + float(rng.uniform(5.0, 50.0))
        out.append((teff, logg, nl, mh, w0, w1))
    return out


def batch_instances(catalog, n: int = 256, seed: int = 3):
    """(atmosphere, catalog, wl0, wl1) per instance of the batched
    configuration, all drawing on one shared line list, and the draws."""
    draws = batch_draws(n, seed)
    inst = [(build_bundle(make_atm(teff, logg, nl, mh)), catalog, w0, w1)
            for teff, logg, nl, mh, w0, w1 in draws]
    return inst, draws
