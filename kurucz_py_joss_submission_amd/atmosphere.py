"""Atmosphere bundles: the per-depth inputs of the hot path.

The on-disk bundle is the reference's ``.npz`` schema (the subset the hot path
reads; synthe_py/io/atmosphere.py:82-506 load_cached, keys written by
tools/convert_atm_to_npz.py:2755-2830):

=================  ===========================  ==============================
key                shape                        meaning
=================  ===========================  ==============================
depth              [D]                          RHOX column mass, g cm^-2
temperature        [D]                          K
gas_pressure       [D]                          dyn cm^-2
electron_density   [D]                          cm^-3
mass_density       [D]                          g cm^-3
turbulent_velocity [D]                          cm s^-1
hckt               [D]                          hc/kT, cm
txnxn              [D] (optional)               vdW perturber density
xnf_h, xnf_he1,    [D]                          for TXNXN when absent
xnf_h2
population_per_ion [D][6][NELEM]                N_ion / U_ion, cm^-3
doppler_per_ion    [D][6][NELEM]                Doppler width / c
wledge             [E]                          continuum edges, nm
half_edge          [E-1] (optional)
delta_edge         [E-1] (optional)
cont_abs_coeff     [D][E-1][3]                  log10 kappa_abs at edge nodes
cont_scat_coeff    [D][E-1][3]                  log10 sigma at edge nodes
cont_coeff_log10   scalar bool
=================  ===========================  ==============================
"""

from __future__ import annotations

from dataclasses import dataclass
from pathlib import Path
from typing import Dict, Optional

import math

import numpy as np


@dataclass
class AtmosphereModel:
    """Hot-path subset of the reference AtmosphereModel (io/atmosphere.py:16-79)."""

    depth: np.ndarray
    temperature: np.ndarray
    gas_pressure: np.ndarray
    electron_density: np.ndarray
    mass_density: np.ndarray
    turbulent_velocity: np.ndarray
    hckt: Optional[np.ndarray] = None
    txnxn: Optional[np.ndarray] = None
    xnf_h: Optional[np.ndarray] = None
    xnf_he1: Optional[np.ndarray] = None
    xnf_h2: Optional[np.ndarray] = None
    xnfph: Optional[np.ndarray] = None  # [D][2] H I and H II populations (hydrogen-line broadening)
    dopph: Optional[np.ndarray] = None  # [D] hydrogen Doppler width / c
    population_per_ion: Optional[np.ndarray] = None
    doppler_per_ion: Optional[np.ndarray] = None
    continuum_wledge: Optional[np.ndarray] = None
    continuum_half_edge: Optional[np.ndarray] = None
    continuum_delta_edge: Optional[np.ndarray] = None
    continuum_abs_coeff: Optional[np.ndarray] = None
    continuum_scat_coeff: Optional[np.ndarray] = None
    continuum_coeff_log10: bool = True
    metadata: Optional[Dict[str, float]] = None
    # NLTE departure coefficients of the BFUDGE column (io/atmosphere.py:56-60), [D] or [D][k]
    bhyd: Optional[np.ndarray] = None
    bc1: Optional[np.ndarray] = None
    bc2: Optional[np.ndarray] = None
    bsi1: Optional[np.ndarray] = None
    bsi2: Optional[np.ndarray] = None

    @property
    def layers(self) -> int:
        return int(self.temperature.size)

    def hckt_or_default(self) -> np.ndarray:
        return depth_hckt(self)

    def txnxn_or_default(self) -> np.ndarray:
        return depth_txnxn(self)

    def to_npz_dict(self) -> Dict[str, np.ndarray]:
        d = {
            "depth": self.depth, "temperature": self.temperature,
            "gas_pressure": self.gas_pressure, "electron_density": self.electron_density,
            "mass_density": self.mass_density, "turbulent_velocity": self.turbulent_velocity,
        }
        opt = {
            "hckt": self.hckt, "txnxn": self.txnxn, "xnf_h": self.xnf_h,
            "xnf_he1": self.xnf_he1, "xnf_h2": self.xnf_h2, "xnfph": self.xnfph, "dopph": self.dopph,
            "population_per_ion": self.population_per_ion,
            "doppler_per_ion": self.doppler_per_ion, "wledge": self.continuum_wledge,
            "half_edge": self.continuum_half_edge, "delta_edge": self.continuum_delta_edge,
            "cont_abs_coeff": self.continuum_abs_coeff,
            "cont_scat_coeff": self.continuum_scat_coeff,
        }
        d.update({k: v for k, v in opt.items() if v is not None})
        d["cont_coeff_log10"] = np.array(bool(self.continuum_coeff_log10))
        for k, v in (self.metadata or {}).items():
            d[f"meta_{k}"] = np.array(v)
        return d

    def save(self, path: Path) -> None:
        np.savez(path, **self.to_npz_dict())


def _opt(atm, name):
    v = getattr(atm, name, None)
    return None if v is None else np.asarray(v, dtype=np.float64)


def depth_hckt(atm) -> np.ndarray:
    """HCKT per depth of any atmosphere with the reference's fields: the
    bundle's value where it has one, else 1.4388 / max(T, 1) (populations.py:162-171)."""
    temp = np.maximum(np.asarray(atm.temperature, dtype=np.float64), 1.0)
    hckt = _opt(atm, "hckt")
    out = 1.4388 / temp
    if hckt is not None:
        k = min(hckt.size, temp.size)
        out[:k] = hckt[:k]
    return out


def depth_txnxn(atm) -> np.ndarray:
    """TXNXN per depth (populations.py:190-199): the bundle's column, else
    (xnf_h + 0.42 xnf_he1 + 0.85 xnf_h2) (T / 10^4)^0.3."""
    tx = _opt(atm, "txnxn")
    if tx is not None:
        return tx
    temp = np.asarray(atm.temperature, dtype=np.float64)
    cols = [_opt(atm, n) for n in ("xnf_h", "xnf_he1", "xnf_h2")]
    out = np.empty(temp.size, dtype=np.float64)
    for i in range(temp.size):  # scalar arithmetic, as the reference's per-depth loop
        t = max(float(temp[i]), 1.0)
        h, he, h2 = (float(c[i]) if c is not None else 0.0 for c in cols)
        out[i] = (h + 0.42 * he + 0.85 * h2) * (t / 10_000.0) ** 0.3
    return out


def depth_rho(atm) -> np.ndarray:
    """Mass density, ones when absent (physics/line_opacity.py:794-797)."""
    v = _opt(atm, "mass_density")
    return np.ones(atm.temperature.size) if v is None else v


def depth_xne(atm) -> np.ndarray:
    """Electron density, zeros when absent (physics/line_opacity.py:798-801)."""
    v = _opt(atm, "electron_density")
    return np.zeros(atm.temperature.size) if v is None else v



def depth_doppler_scale(atm, microturb_kms: float = 0.0, mass_amu: float = 56.0) -> np.ndarray:
    """Doppler width / line wavelength per depth as physics/populations.py:172-190
    computes it for a line of mass_amu (the engine passes no elements, so every
    line takes the default 56): sqrt(v_turb^2 + 2kT/m) / c with v_turb the
    model's turbulence combined with the synthesis microturbulence."""
    temp = np.asarray(atm.temperature, dtype=np.float64)
    vt = _opt(atm, "turbulent_velocity")
    out = np.empty(temp.size, dtype=np.float64)
    mass = np.array([mass_amu], dtype=np.float64)
    for i in range(temp.size):
        t = max(float(temp[i]), 1.0)
        thermal = np.sqrt(2.0 * 1.380649e-16 * t / (mass * 1.66054e-24)) / 2.99792458e10
        v_model = (float(vt[i]) if vt is not None and vt.size > i else 0.0) / 1e5 / 299_792.458
        micro = microturb_kms / 299_792.458 if microturb_kms > 0 else 0.0
        turb = math.sqrt(v_model * v_model + micro * micro)
        out[i] = float(np.sqrt(turb ** 2 + thermal ** 2)[0])
    return out


def merge_energies(atm):
    """(EMERGE, EMERGE_H) per depth, cm^-1 (engine/opacity.py:2368-2372): the
    Inglis-Teller last series member from the electron density."""
    ne = np.maximum(np.asarray(atm.electron_density, dtype=np.float64), 1e-40)
    inglis = 1600.0 / np.power(ne, 2.0 / 15.0)
    nmerge = np.maximum(inglis - 1.5, 1.0)
    return 109737.312 / np.maximum(nmerge ** 2, 1e-12), 109677.576 / np.maximum(nmerge ** 2, 1e-12)


def _xnfph(xnfph, qxnfpel):
    """XNFPH [D][2] (H I, H II) as io/atmosphere.py:157-161 reads it: the
    archive's own, else the first two columns of the legacy QXNFPEL."""
    if xnfph is not None:
        return np.asarray(xnfph, np.float64)
    if qxnfpel is not None and qxnfpel.ndim == 2 and qxnfpel.shape[1] >= 2:
        return np.ascontiguousarray(qxnfpel[:, :2], np.float64)
    return None


def _dopph(dopph, qdopple):
    """DOPPH [D] as io/atmosphere.py:163-167 reads it (QDOPPLE's first column)."""
    if dopph is not None:
        return np.asarray(dopph, np.float64)
    if qdopple is not None and qdopple.ndim == 2:
        return np.ascontiguousarray(qdopple[:, 0], np.float64)
    return None

def load_cached(path: Path) -> AtmosphereModel:
    """Load a bundle ``.npz`` (io/atmosphere.py:82-506, hot-path keys only)."""
    with np.load(path, allow_pickle=False) as data:
        def pick(*names, required=True):
            for n in names:
                if n in data:
                    return np.asarray(data[n])
            if required:
                raise KeyError(f"Required atmosphere field missing: {' / '.join(names)}")
            return None

        meta = {k[5:]: data[k].item() for k in data.files if k.startswith("meta_")}
        return AtmosphereModel(
            depth=pick("depth", "rhox", "qrhox").astype(np.float64),
            temperature=pick("temperature", "qt").astype(np.float64),
            gas_pressure=pick("gas_pressure", "qp").astype(np.float64),
            electron_density=pick("electron_density", "qxne").astype(np.float64),
            mass_density=pick("mass_density", "qrho").astype(np.float64),
            turbulent_velocity=pick("turbulent_velocity", "qvturb").astype(np.float64),
            hckt=pick("hckt", "qhckt", required=False),
            txnxn=pick("txnxn", required=False),
            xnf_h=pick("xnf_h", "qxnf_h", required=False),
            xnf_he1=pick("xnf_he1", required=False),
            xnf_h2=pick("xnf_h2", "qxnf_h2", required=False),
            xnfph=_xnfph(pick("xnfph", required=False), pick("qxnfpel", required=False)),
            dopph=_dopph(pick("dopph", required=False), pick("qdopple", required=False)),
            population_per_ion=pick("population_per_ion", required=False),
            doppler_per_ion=pick("doppler_per_ion", required=False),
            continuum_wledge=pick("wledge", required=False),
            continuum_half_edge=pick("half_edge", required=False),
            continuum_delta_edge=pick("delta_edge", required=False),
            continuum_abs_coeff=pick("cont_abs_coeff", required=False),
            continuum_scat_coeff=pick("cont_scat_coeff", required=False),
            continuum_coeff_log10=bool(data["cont_coeff_log10"]) if "cont_coeff_log10" in data else False,
            metadata=meta,
            bhyd=pick("bhyd", required=False),
            bc1=pick("bc1", required=False),
            bc2=pick("bc2", required=False),
            bsi1=pick("bsi1", required=False),
            bsi2=pick("bsi2", required=False),
        )
