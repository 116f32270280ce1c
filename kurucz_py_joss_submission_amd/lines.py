"""Atomic line lists: GFALL fixed-width I/O and the line SoA the kernels read.

Mirrors the reference's line semantics (synthe_py/io/lines/atomic.py):

* wavelength is recomputed from the level energies, lambda = 1e7/|EP - E| nm
  (+ DWLISO), and used for grid indexing and CGF (atomic.py:437-457);
* gf = 10**(log gf + X1 + X2) (atomic.py:381-383, LineCatalog.from_records 275);
* E_low = min(|E|, |EP|) in cm^-1 (atomic.py:389-391);
* gamma = 10**G / (12.5664 * nu), nu = 2.99792458e17/lambda (atomic.py:495-504),
  with G == 0 meaning "use the default" (radiative default 2.223e13/lambda^2,
  atomic.py:466-469);
* ion stage = round(frac(CODE)*100) + 1 (atomic.py:508-511);
* the DELLIM window filter (atomic.py:712-816).

The parser is vectorised over fixed-width records (one numpy pass instead of a
Python loop per line); every arithmetic step is the reference's, in the same
order.  Stark/van der Waals defaults need the POTION ionisation table and are
not implemented: records with GS == 0 or GW == 0 raise.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from pathlib import Path
from typing import Optional

import numpy as np

from .constants import (
    C_LIGHT_NM,
    CGF_CONSTANT,
    ELEMENT_SYMBOLS,
    ELEMENT_Z,
)

# DELLIM margins in nm (atomic.py:693)
_DELLIM = np.array([100.0, 30.0, 10.0, 3.0, 1.0, 0.3, 0.1], dtype=np.float64)


@dataclass
class LineCatalog:
    """Structure-of-arrays line catalog (atomic.py:202-296 LineCatalog fields)."""

    wavelength: np.ndarray
    index_wavelength: np.ndarray
    log_gf: np.ndarray
    gf: np.ndarray
    excitation_energy: np.ndarray
    gamma_rad: np.ndarray
    gamma_stark: np.ndarray
    gamma_vdw: np.ndarray
    atomic_number: np.ndarray  # int32 Z (0 if unknown)
    ion_stages: np.ndarray  # int16
    line_types: np.ndarray  # int8
    n_lower: np.ndarray = field(default=None)
    n_upper: np.ndarray = field(default=None)
    code: np.ndarray = field(default=None)
    line_size: np.ndarray = field(default=None)

    def __len__(self) -> int:
        return int(self.wavelength.size)

    @property
    def elements(self) -> np.ndarray:
        return np.array([ELEMENT_SYMBOLS[z] if 0 < z < len(ELEMENT_SYMBOLS) else f"Z{z}"
                         for z in self.atomic_number], dtype=object)

    def subset(self, idx: np.ndarray) -> "LineCatalog":
        kw = {}
        for name in self.__dataclass_fields__:
            v = getattr(self, name)
            kw[name] = None if v is None else v[idx]
        return LineCatalog(**kw)

    def save_npz(self, path: Path) -> None:
        np.savez(path, **{k: getattr(self, k) for k in self.__dataclass_fields__
                          if getattr(self, k) is not None})

    @classmethod
    def load_npz(cls, path: Path) -> "LineCatalog":
        with np.load(path, allow_pickle=False) as z:
            return cls(**{k: z[k] for k in z.files})


def _field(block: np.ndarray, a: int, b: int) -> np.ndarray:
    """Columns [a, b) of a (n, width) uint8 record block as stripped strings."""
    raw = np.ascontiguousarray(block[:, a:b]).view(f"S{b - a}").ravel()
    return np.char.strip(raw)


def _to_float(col: np.ndarray, default: float = 0.0) -> np.ndarray:
    out = np.full(col.shape, default, dtype=np.float64)
    nz = col != b""
    if np.any(nz):
        out[nz] = col[nz].astype(np.float64)
    return out


def load_catalog(path: Path) -> LineCatalog:
    """Parse a GFALL fixed-width file (atomic.py:299-610, fixed-width path).

    Records must be at least 98 characters wide; comment and blank lines are
    skipped.  Raises ValueError for records the hot path cannot represent.
    """
    text = Path(path).read_bytes().splitlines()
    recs = [ln for ln in text if ln.strip() and not ln.strip().startswith(b"#")]
    if not recs:
        return _empty_catalog()
    width = max(len(r) for r in recs)
    if min(len(r) for r in recs) < 98:
        raise ValueError("GFALL records shorter than 98 columns are not supported")
    width = max(width, 160)
    block = np.zeros((len(recs), width), dtype=np.uint8) + ord(" ")
    for i, r in enumerate(recs):
        block[i, : len(r)] = np.frombuffer(r, dtype=np.uint8)
    lens = np.array([len(r) for r in recs])

    wl_stored = _to_float(_field(block, 0, 11))
    log_gf_raw = _to_float(_field(block, 11, 18))
    code = _to_float(_field(block, 18, 24))
    e_val = _to_float(_field(block, 24, 36))
    ep_val = _to_float(_field(block, 52, 64))
    gr_log = _to_float(_field(block, 80, 86))
    gs_log = _to_float(_field(block, 86, 92))
    gw_log = _to_float(_field(block, 92, 98))
    x1 = np.where(lens >= 115, _to_float(_field(block, 109, 115)), 0.0)
    x2 = np.where(lens >= 124, _to_float(_field(block, 118, 124)), 0.0)
    nblo = np.where(lens >= 106, _to_float(_field(block, 102, 104)), 0.0).astype(np.int16)
    nbup = np.where(lens >= 106, _to_float(_field(block, 104, 106)), 0.0).astype(np.int16)
    iso1 = np.where(lens >= 109, _to_float(_field(block, 106, 109)), 0.0).astype(np.int32)
    cother1 = _field(block, 124, 134)
    isoshift = np.where(lens >= 160, _to_float(_field(block, 154, 160)), 0.0)
    cother2 = np.ascontiguousarray(block[:, 134:144]).view("S10").ravel()
    line_size_chr = np.ascontiguousarray(block[:, 140:141]).view("S1").ravel()
    auto_tag = np.char.strip(np.ascontiguousarray(block[:, 141:144]).view("S3").ravel())

    if np.any(cother1 != b""):
        # hyperfine shifts: ishift/ishiftp in mK (atomic.py:421-432)
        ishift = _to_float(_field(block, 124, 129))
        ishiftp = _to_float(_field(block, 129, 134))
        eshift = ishift * 1.0e-3
        eshiftp = ishiftp * 1.0e-3
    else:
        eshift = np.zeros_like(e_val)
        eshiftp = np.zeros_like(e_val)

    log_gf = log_gf_raw + (x1 + x2)
    excitation_cm = np.minimum(np.abs(e_val), np.abs(ep_val))

    wavelength = wl_stored.copy()
    energy_diff = np.abs((np.abs(ep_val) + eshiftp) - (np.abs(e_val) + eshift))
    pos = energy_diff > 0.0
    dwliso = isoshift * 1.0e-4
    wavelength[pos] = 1.0e7 / energy_diff[pos] + dwliso[pos]

    if np.any(gs_log == 0.0) or np.any(gw_log == 0.0):
        raise ValueError("Stark / van der Waals defaults (GS or GW == 0) need the POTION "
                         "table and are outside the hot path's line lists")
    gamma_rad = np.array([10.0 ** v if v != 0.0 else 0.0 for v in gr_log], dtype=np.float64)
    gamma_stark = np.array([10.0 ** v for v in gs_log], dtype=np.float64)
    gamma_vdw = np.array([10.0 ** v for v in gw_log], dtype=np.float64)
    zero_r = gamma_rad == 0.0
    gamma_rad[zero_r] = 2.223e13 / (wavelength[zero_r] ** 2)
    frelin = C_LIGHT_NM / np.maximum(wavelength, 1e-12)
    gamma_rad = gamma_rad / (12.5664 * frelin)
    gamma_stark = gamma_stark / (12.5664 * frelin)
    gamma_vdw = gamma_vdw / (12.5664 * frelin)

    nelem = (code + 1e-6).astype(np.int64)
    frac = code - nelem
    ion_stage = np.where(frac > 1e-6, np.round(frac * 100.0).astype(np.int64) + 1, 1).astype(np.int16)
    z = np.where((nelem > 0) & (nelem < len(ELEMENT_SYMBOLS)), nelem, 0).astype(np.int32)

    line_size = np.array([int(c) if c.isdigit() else 0 for c in line_size_chr.tolist()],
                         dtype=np.int16)
    line_type = np.zeros(len(recs), dtype=np.int8)
    tags = auto_tag.tolist()
    for i, t in enumerate(tags):
        if t == b"AUT":
            line_type[i] = 1
        elif t == b"COR":
            line_type[i] = 2
        elif t == b"PRD":
            line_type[i] = 3
    plain = line_type == 0
    line_type[plain & (nelem == 1) & (ion_stage == 1) & (iso1 == 2)] = -2
    line_type[plain & (nelem == 1) & (ion_stage == 1) & (iso1 != 2)] = -1
    line_type[plain & (nelem == 2) & (ion_stage == 1) & (iso1 == 3)] = -4
    line_type[plain & (nelem == 2) & (ion_stage == 1) & (iso1 != 3)] = -3
    line_type[plain & (nelem == 2) & (ion_stage == 2)] = -6
    del cother2

    gf = np.power(10.0, log_gf)
    return LineCatalog(
        wavelength=wavelength,
        index_wavelength=wavelength.copy(),
        log_gf=log_gf,
        gf=gf,
        excitation_energy=excitation_cm,
        gamma_rad=gamma_rad,
        gamma_stark=gamma_stark,
        gamma_vdw=gamma_vdw,
        atomic_number=z,
        ion_stages=ion_stage,
        line_types=line_type,
        n_lower=nblo,
        n_upper=nbup,
        code=code,
        line_size=line_size,
    )


def _empty_catalog() -> LineCatalog:
    e = np.zeros(0, dtype=np.float64)
    return LineCatalog(e, e, e, e, e, e, e, e, np.zeros(0, np.int32), np.zeros(0, np.int16),
                       np.zeros(0, np.int8), np.zeros(0, np.int16), np.zeros(0, np.int16), e,
                       np.zeros(0, np.int16))


def line_margin(line_type: np.ndarray, atomic_number: np.ndarray, wl_min: float,
                line_size: Optional[np.ndarray] = None, code: Optional[np.ndarray] = None) -> np.ndarray:
    """DELLIM margin per line in nm (atomic.py:697-740, _get_line_margin)."""
    delfactor = 1.0 if wl_min <= 500.0 else wl_min / 500.0
    n = line_type.size
    ls = np.zeros(n, dtype=np.int64) if line_size is None else np.maximum(line_size.astype(np.int64), 0)
    lim = np.minimum(8 - ls, 7)
    is_h = (line_type == -1) | (line_type == -2) | (atomic_number == 1)
    if code is not None:
        is_h |= np.abs(code - 1.0) < 1.0e-6
    lim = np.where(is_h, 1, lim)
    return _DELLIM[lim - 1] * delfactor


def filter_by_range(catalog: LineCatalog, wl_min: float, wl_max: float) -> LineCatalog:
    """Keep lines inside [wl_min - margin, wl_max + margin] (atomic.py:743-816)."""
    if len(catalog) == 0:
        return catalog
    margin = line_margin(catalog.line_types, catalog.atomic_number, wl_min, catalog.line_size,
                         catalog.code)
    wl = catalog.wavelength
    keep = np.nonzero((wl >= wl_min - margin) & (wl <= wl_max + margin))[0]
    return catalog.subset(keep)


def write_gfall(path: Path, wl_nm: np.ndarray, log_gf: np.ndarray, code: np.ndarray,
                e_low_cm: np.ndarray, gr_log: np.ndarray, gs_log: np.ndarray,
                gw_log: np.ndarray) -> None:
    """Write 160-column GFALL records with EP = E + 1e7/lambda.

    Field layout follows atomic.py:325-345 (F11.4, F7.3, F6.2, F12.3, F5.1,
    1X, A8, A2, F12.3, F5.1, 1X, A8, A2, 3F6.2, A4, 2I2, I3, F6.3, I3, F6.3,
    ...).  The parsed wavelength 1e7/|EP - E| is the canonical one.
    """
    ep = e_low_cm + 1.0e7 / wl_nm
    out = []
    for i in range(wl_nm.size):
        rec = (
            f"{wl_nm[i]:11.4f}{log_gf[i]:7.3f}{code[i]:6.2f}{e_low_cm[i]:12.3f}{0.5:5.1f}"
            f" {'LOWLEV':8s}{'  ':2s}{ep[i]:12.3f}{1.5:5.1f} {'UPLEV':8s}{'  ':2s}"
            f"{gr_log[i]:6.2f}{gs_log[i]:6.2f}{gw_log[i]:6.2f}{'SYN ':4s}{0:2d}{0:2d}"
            f"{0:3d}{0.0:6.3f}{0:3d}{0.0:6.3f}{'':10s}{'':10s}{'':10s}{0:6d}"
        )
        assert len(rec) == 160, len(rec)
        out.append(rec)
    Path(path).write_text("\n".join(out) + "\n")


@dataclass
class LineSoA:
    """Per-line arrays the kernels consume (line_opacity.py:740-784 prep loop).

    process: 1 if the line routes to the plain-line (fort.12) path and its
    element is present in the atmosphere bundle.  elem is Z-1, ion the stage
    (1-based).  center_t is the TRANSP/centre index (engine/opacity.py:550-592),
    center_w the wing index (line_opacity.py:1030-1059).
    """

    wl: np.ndarray
    gf: np.ndarray
    cgf: np.ndarray
    gamma_rad: np.ndarray
    gamma_stark: np.ndarray
    gamma_vdw: np.ndarray
    elo: np.ndarray
    elem: np.ndarray
    ion: np.ndarray
    ltype: np.ndarray
    process: np.ndarray
    center_t: np.ndarray
    center_w: np.ndarray

    def __len__(self) -> int:
        return int(self.wl.size)


def nearest_grid_indices(grid: np.ndarray, values: np.ndarray) -> np.ndarray:
    """Centre index by log rounding (engine/opacity.py:550-592)."""
    if len(grid) < 2:
        return np.zeros(len(values), dtype=np.int64)
    ratio = grid[1] / grid[0]
    ratiolg = np.log(ratio)
    ix_start = int(np.log(grid[0]) / ratiolg + 0.5)
    with np.errstate(divide="ignore", invalid="ignore"):
        ixwl = (np.log(values) / ratiolg + 0.5).astype(np.int64)
        idx = ixwl - ix_start
    idx[values < grid[0]] = -1
    idx[values > grid[-1]] = grid.size
    return idx


def wing_center_indices(grid: np.ndarray, values: np.ndarray,
                        origin: Optional[float] = None) -> np.ndarray:
    """Unclamped wing centre index by rint (line_opacity.py:1030-1059)."""
    if len(grid) < 2:
        return np.zeros(len(values), dtype=np.int64)
    ratio = grid[1] / grid[0]
    ratiolg = np.log(ratio)
    start_val = grid[0] if origin is None else origin
    ix_floor = int(np.floor(np.log(start_val) / ratiolg))
    wbegin = np.exp(ix_floor * ratiolg)
    if wbegin < start_val:
        ix_floor += 1
        wbegin = np.exp(ix_floor * ratiolg)
    with np.errstate(divide="ignore", invalid="ignore"):
        ix = np.rint(np.log(values / wbegin) / ratiolg).astype(np.int64)
    if origin is not None:
        ix_floor = int(np.floor(np.log(origin) / ratiolg))
        wb = np.exp(ix_floor * ratiolg)
        if wb < origin:
            ix_floor += 1
            wb = np.exp(ix_floor * ratiolg)
        off = int(np.rint(np.log(grid[0] / wb) / ratiolg))
        ix = ix - off
    return ix


def build_line_soa(catalog: LineCatalog, grid: np.ndarray, n_elements: int,
                   include_h_lines: bool = False, grid_origin: Optional[float] = None) -> LineSoA:
    """Host prep of the per-line arrays (line_opacity.py:740-784, 1030-1059)."""
    lt = catalog.line_types.astype(np.int32)
    nlo = np.abs(catalog.n_lower.astype(np.int64)) if catalog.n_lower is not None else 0
    nup = np.abs(catalog.n_upper.astype(np.int64)) if catalog.n_upper is not None else 0
    nb_sum = nlo + nup
    routes = (lt != 2) & (lt != 1) & (lt <= 3) & (nb_sum == 0)
    z = catalog.atomic_number.astype(np.int64)
    is_h = (z == 1) & (catalog.ion_stages == 1)
    if not include_h_lines:
        routes &= ~((lt == -1) | is_h)
    known = (z > 0) & (z - 1 < n_elements)
    process = routes & known
    iw = catalog.index_wavelength
    wl = np.where(process, iw, 0.0)
    freq_hz = np.where(process, C_LIGHT_NM / np.where(process, iw, 1.0), 1.0)
    cgf = np.where(process, CGF_CONSTANT * catalog.gf / freq_hz, 0.0)
    return LineSoA(
        wl=np.ascontiguousarray(iw, dtype=np.float64),
        gf=np.ascontiguousarray(catalog.gf, dtype=np.float64),
        cgf=np.ascontiguousarray(cgf, dtype=np.float64),
        gamma_rad=np.ascontiguousarray(catalog.gamma_rad, dtype=np.float64),
        gamma_stark=np.ascontiguousarray(catalog.gamma_stark, dtype=np.float64),
        gamma_vdw=np.ascontiguousarray(catalog.gamma_vdw, dtype=np.float64),
        elo=np.ascontiguousarray(catalog.excitation_energy, dtype=np.float64),
        elem=np.ascontiguousarray(np.where(process, z - 1, -1), dtype=np.int32),
        ion=np.ascontiguousarray(catalog.ion_stages, dtype=np.int32),
        ltype=np.ascontiguousarray(lt, dtype=np.int32),
        process=np.ascontiguousarray(process, dtype=np.uint8),
        center_t=np.ascontiguousarray(nearest_grid_indices(grid, iw), dtype=np.int64),
        center_w=np.ascontiguousarray(wing_center_indices(grid, iw, grid_origin),
                                      dtype=np.int64),
    )
