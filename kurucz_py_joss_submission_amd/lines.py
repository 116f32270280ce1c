"""Atomic line lists: GFALL fixed-width I/O and the line SoA the kernels read.

Mirrors the reference's line semantics (synthe_py/io/lines/atomic.py):

* wavelength is recomputed from the level energies, lambda = 1e7/|EP - E| nm
  (+ DWLISO), and used for grid indexing and CGF (atomic.py:437-457);
* gf = 10**(log gf + X1 + X2) (atomic.py:381-383, LineCatalog.from_records 275);
* E_low = min(|E|, |EP|) in cm^-1 (atomic.py:389-391);
* gamma = 10**G / (12.5664 * nu), nu = 2.99792458e17/lambda (atomic.py:495-504),
  with G == 0 meaning "use the default": radiative 2.223e13/lambda^2
  (atomic.py:466-469), Stark and van der Waals from the effective quantum
  numbers of the levels against the POTION ionisation potentials
  (atomic.py:59-199, applied at 476-492);
* ion stage = round(frac(CODE)*100) + 1 (atomic.py:508-511);
* records that are not fixed-width GFALL go to the whitespace reader
  (atomic.py:594-672);
* the DELLIM window filter (atomic.py:712-816).

The parser is vectorised over fixed-width records (one numpy pass instead of a
Python loop per line); every arithmetic step is the reference's, in the same
order, with the same routine (numpy or the C library) where the two round
differently.
"""

from __future__ import annotations

import math
import os
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
# the GFALL reader names elements up to uranium; heavier codes become "Z<n>"
# and no plain-line kernel takes them (io/lines/tfort.py:14-108)
PARSER_MAX_Z = 92

_POTION: Optional[np.ndarray] = None


def potion() -> np.ndarray:
    """Ionisation potentials in cm^-1 by POTION slot (synthe_py/data/fortran_data.npz)."""
    global _POTION
    if _POTION is None:
        with np.load(Path(__file__).resolve().parent / "data" / "kurucz_tables.npz", allow_pickle=False) as z:
            _POTION = np.asarray(z["potion"], dtype=np.float64)
    return _POTION


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
    # rgfall metadata the fort.19 records are built from (atomic.py:202-232):
    # ISO1/ISO2, the J values, LABELP starting "CONTINUU" and the raw log
    # GR/GS/GW columns (0 where blank)
    iso1: np.ndarray = field(default=None)
    iso2: np.ndarray = field(default=None)
    xj: np.ndarray = field(default=None)
    xjp: np.ndarray = field(default=None)
    continuum_label: np.ndarray = field(default=None)  # uint8
    gr_log: np.ndarray = field(default=None)
    gs_log: np.ndarray = field(default=None)
    gw_log: np.ndarray = field(default=None)

    def __len__(self) -> int:
        return int(self.wavelength.size)

    @property
    def elements(self) -> np.ndarray:
        """Element symbols as the reference's records carry them ("Z<n>" past uranium)."""
        out = []
        for i, z in enumerate(self.atomic_number.tolist()):
            if 0 < z < len(ELEMENT_SYMBOLS):
                out.append(ELEMENT_SYMBOLS[z])
            else:
                n = int(self.code[i] + 1e-6) if self.code is not None else z
                out.append(f"Z{n}")
        return np.array(out, dtype=object)

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


def _floats(col: np.ndarray, default: float = 0.0):
    """float() of each stripped field (blank -> default) and where it parsed."""
    out = np.full(col.shape, default, dtype=np.float64)
    ok = np.ones(col.shape, dtype=bool)
    nz = col != b""
    if np.any(nz):
        try:
            out[nz] = col[nz].astype(np.float64)
        except ValueError:
            for i in np.nonzero(nz)[0]:
                try:
                    out[i] = float(col[i])
                except ValueError:
                    ok[i] = False
    return out, ok


def _ints(col: np.ndarray, default: int = 0) -> np.ndarray:
    """int() of each field, default where blank or not an integer."""
    uniq, inv = np.unique(col, return_inverse=True)
    vals = np.empty(uniq.size, dtype=np.int64)
    for k, u in enumerate(uniq.tolist()):
        try:
            vals[k] = int(u) if u.strip() else default
        except ValueError:
            vals[k] = default
    return vals[inv.reshape(col.shape)]


def _code_parts(code: np.ndarray):
    nelem = (code + 1e-6).astype(np.int64)
    icharge = ((code - nelem) * 100.0 + 0.1).astype(np.int64)
    return nelem, icharge, icharge + 1


def _potion_at(nelem: np.ndarray, icharge: np.ndarray):
    """POTION value of each (element, charge) and whether the slot exists (atomic.py:40-56)."""
    slot = np.where(nelem <= 30, nelem * (nelem + 1) // 2 + icharge, nelem * 5 + 341 + icharge) - 1
    tab = potion()
    inside = (slot >= 0) & (slot < tab.size)
    return np.where(inside, tab[np.clip(slot, 0, tab.size - 1)], 0.0), inside


def stark_default_log(code: np.ndarray, e_lower: np.ndarray, e_upper: np.ndarray) -> np.ndarray:
    """log10 GAMMAS of records without one (atomic.py:59-106): an effective
    principal quantum number of the upper level from its distance to the
    ionisation limit, n*^2 = R Z_eff^2 / (IP - E_up) (25 when unbound),
    GAMMAS = 1e-8 n*^5; molecules (CODE >= 100) take 1e-5."""
    code = np.asarray(code, np.float64)
    nelem, icharge, zeff = _code_parts(code)
    eup = np.maximum(np.abs(e_lower), np.abs(e_upper))
    ip, inside = _potion_at(nelem, icharge)
    gap = ip - eup
    n2 = np.full(code.shape, 25.0)
    bound = inside & (gap > 0)
    n2[bound] = 109737.31 * zeff[bound] * zeff[bound] / gap[bound]
    out = np.empty(code.shape)
    for i in range(code.size):  # the reference's scalar math, for the same roundings
        out[i] = -5.0 if code[i] >= 100.0 else math.log10(1.0e-8 * n2[i] * n2[i] * math.sqrt(n2[i]))
    return out


def vdw_default_log(code: np.ndarray, e_lower: np.ndarray, e_upper: np.ndarray) -> np.ndarray:
    """log10 GAMMAW of records without one (atomic.py:109-199): Unsold's
    approximation 4.5e-9 (<r^2>_up - <r^2>_lo)^0.4 with hydrogenic
    <r^2> = 2.5 (n*^2 / Z_eff)^2, n*^2 capped at 1000; the transition-metal
    sequences 21-28 take (45 - NSEQ)/Z_eff; molecules take 1e-7 / Z_eff."""
    code = np.asarray(code, np.float64)
    nelem, icharge, zeff = _code_parts(code)
    elo = np.minimum(np.abs(e_lower), np.abs(e_upper))
    eup = np.maximum(np.abs(e_lower), np.abs(e_upper))
    ip, _ = _potion_at(nelem, icharge)
    out = np.empty(code.shape)
    for i in range(code.size):
        z = int(zeff[i])
        if code[i] >= 100.0:
            out[i] = math.log10(1.0e-7 / z)
            continue
        r2 = []
        for level in (eup[i], elo[i]):
            n2 = 25.0
            if ip[i] > 0 and ip[i] - level > 0:
                n2 = 109737.31 * z * z / (ip[i] - level)
            n2 = min(n2, 1000.0)
            r2.append(2.5 * (n2 / z) ** 2)
        r_up, r_lo = r2
        nseq = int(nelem[i]) - z + 1
        if 20 < nseq < 29:
            r_up, r_lo = (45.0 - nseq) / z, 0.0
        if r_up < r_lo:
            r_up = 2.0 * r_lo
        d = r_up - r_lo
        out[i] = math.log10(4.5e-9 * (d ** 0.4) if d > 0 else 1.0e-9)
    return out


def _pow10_or_zero(log_values: np.ndarray) -> np.ndarray:
    # scalar pow, as the reference exponentiates each record's G
    return np.array([10.0 ** v if v != 0.0 else 0.0 for v in log_values.tolist()], dtype=np.float64)


def _whitespace_record(line: str):
    """A record in the whitespace format (atomic.py:594-672), or None."""
    parts = line.split()
    if len(parts) < 7:
        return None
    try:
        wl = float(parts[0])
        sym = parts[1]
        try:
            stage = int(float(parts[2]))
        except (ValueError, IndexError):
            return None
        vals = [float(parts[k]) for k in (3, 4, 5, 6)]
        gw = float(parts[7]) if len(parts) > 7 else 0.0
        nlo = nup = 0
        if len(parts) > 9:
            try:
                nlo = int(float(parts[8]))
                nup = int(float(parts[9]))
            except ValueError:
                pass
        lt = {("H", 1): -1, ("He", 1): -3, ("He", 2): -6}.get((sym, stage), 0)
        z = ELEMENT_Z.get(sym.strip().upper().replace(" ", ""), 0)
        return dict(wl=wl, z=z, stage=stage, log_gf=vals[0], elo=vals[1],
                    gr=vals[2], gs=vals[3], gw=gw, lt=lt, nlo=nlo, nup=nup)
    except (ValueError, IndexError):
        return None


def load_catalog(path: Path) -> LineCatalog:
    """Parse a GFALL file (atomic.py:299-686).

    Fixed-width records (98 columns or more) are parsed column-wise; a record
    that is shorter, or whose fixed-width fields do not parse, is read in the
    whitespace format instead, and skipped if that fails too, as the
    reference does.  Comment and blank lines are skipped.
    """
    with Path(path).open("r", encoding="ascii", errors="ignore") as fh:
        text = fh.read().splitlines()
    recs = [ln for ln in text if ln.strip() and not ln.strip().startswith("#")]
    if not recs:
        return _empty_catalog()
    n_rec = len(recs)
    lens = np.array([len(r) for r in recs])
    width = max(int(lens.max()), 160)
    fixed = lens >= 98
    block = np.full((n_rec, width), ord(" "), dtype=np.uint8)
    for i, r in enumerate(recs):
        block[i, : len(r)] = np.frombuffer(r.encode("ascii"), dtype=np.uint8)

    wl_stored, ok0 = _floats(_field(block, 0, 11))
    log_gf_raw, ok1 = _floats(_field(block, 11, 18))
    code, ok2 = _floats(_field(block, 18, 24))
    e_col = _field(block, 24, 36)
    e_val, ok3 = _floats(e_col)
    ep_val, ok4 = _floats(_field(block, 52, 64))
    gr_log, ok5 = _floats(_field(block, 80, 86))
    gs_log, ok6 = _floats(_field(block, 86, 92))
    gw_log, ok7 = _floats(_field(block, 92, 98))
    xj, ok8 = _floats(_field(block, 36, 41))  # J values: parsed, a bad one fails the record
    xjp, ok9 = _floats(_field(block, 64, 69))
    good = fixed & ok0 & ok1 & ok2 & ok3 & ok4 & ok5 & ok6 & ok7 & ok8 & ok9
    good &= (_field(block, 0, 11) != b"") & (_field(block, 11, 18) != b"") & (_field(block, 18, 24) != b"")
    good &= e_col != b""

    x1, okx1 = _floats(_field(block, 109, 115))
    x2, okx2 = _floats(_field(block, 118, 124))
    x1 = np.where((lens >= 115) & okx1, x1, 0.0)
    x2 = np.where((lens >= 124) & okx2, x2, 0.0)
    nblo_s, nbup_s = _field(block, 102, 104), _field(block, 104, 106)
    nblo = np.where(lens >= 106, _ints(nblo_s), 0)
    nbup = np.where(lens >= 106, _ints(nbup_s), 0)
    # int() of NBUP failing after NBLO parsed keeps NBLO (atomic.py:524-531)
    bad_lo = (nblo_s != b"") & (_ints(nblo_s, -99999) == -99999)
    nbup = np.where(bad_lo, 0, nbup)
    nblo = nblo.astype(np.int16)
    nbup = nbup.astype(np.int16)
    iso1 = np.where(lens >= 109, _ints(_field(block, 106, 109)), 0)
    iso2 = np.where(lens >= 119, _ints(_field(block, 116, 119)), 0)
    labelp = np.where(lens >= 78, _field(block, 70, 78), b"")
    continuum_label = np.char.startswith(np.char.upper(labelp), b"CONTINUU").astype(np.uint8)
    isoshift = np.where(lens >= 160, _ints(_field(block, 154, 160)), 0)
    hyper = np.ascontiguousarray(block[:, 124:134]).view("S10").ravel()
    has_hyper = (lens >= 134) & (np.char.strip(hyper) != b"")
    eshift = np.zeros(n_rec)
    eshiftp = np.zeros(n_rec)
    if np.any(has_hyper):
        # ISHIFT / ISHIFTP in mK; a field that is not an integer zeroes both (atomic.py:421-432)
        s1 = _ints(np.ascontiguousarray(block[:, 124:129]).view("S5").ravel(), -99999)
        s2 = _ints(np.ascontiguousarray(block[:, 129:134]).view("S5").ravel(), -99999)
        both = has_hyper & (s1 != -99999) & (s2 != -99999)
        eshift = np.where(both, s1 * 1.0e-3, 0.0)
        eshiftp = np.where(both, s2 * 1.0e-3, 0.0)
    tail_ok = (lens >= 144) & (np.char.strip(np.ascontiguousarray(block[:, 134:144]).view("S10").ravel()) != b"")
    line_size_chr = np.ascontiguousarray(block[:, 140:141]).view("S1").ravel()
    auto_tag = np.char.strip(np.ascontiguousarray(block[:, 141:144]).view("S3").ravel())

    if os.environ.get("PY_APPLY_ISO_CORR", "1") != "0":
        log_gf = log_gf_raw + (x1 + x2)
    else:
        log_gf = log_gf_raw + 0.0
    excitation_cm = np.minimum(np.abs(e_val), np.abs(ep_val))
    excitation_up = np.maximum(np.abs(e_val), np.abs(ep_val))

    wavelength = wl_stored.copy()
    energy_diff = np.abs((np.abs(ep_val) + eshiftp) - (np.abs(e_val) + eshift))
    pos = energy_diff > 0.0
    dwliso = isoshift * 1.0e-4
    wavelength[pos] = 1.0e7 / energy_diff[pos] + dwliso[pos]

    gamma_rad = _pow10_or_zero(gr_log)
    gamma_stark = _pow10_or_zero(gs_log)
    gamma_vdw = _pow10_or_zero(gw_log)
    zero_r = gamma_rad == 0.0
    with np.errstate(divide="ignore"):  # rows that failed to parse are dropped below
        gamma_rad[zero_r] = 2.223e13 / (wavelength[zero_r] ** 2)
    for gam, default in ((gamma_stark, stark_default_log), (gamma_vdw, vdw_default_log)):
        zero = np.nonzero(good & (gam == 0.0))[0]
        if zero.size:
            gam[zero] = _pow10_or_zero(default(code[zero], excitation_cm[zero], excitation_up[zero]))
    frelin = C_LIGHT_NM / np.maximum(wavelength, 1e-12)
    gamma_rad = gamma_rad / (12.5664 * frelin)
    gamma_stark = gamma_stark / (12.5664 * frelin)
    gamma_vdw = gamma_vdw / (12.5664 * frelin)

    nelem = (code + 1e-6).astype(np.int64)
    frac = code - nelem
    ion_stage = np.where(frac > 1e-6, np.round(frac * 100.0).astype(np.int64) + 1, 1).astype(np.int16)
    z = np.where((nelem > 0) & (nelem <= PARSER_MAX_Z), nelem, 0).astype(np.int32)

    line_size = np.array([int(c) if c.isdigit() else 0 for c in line_size_chr.tolist()], dtype=np.int16)
    line_size = np.where(tail_ok, line_size, 0).astype(np.int16)
    line_type = np.zeros(n_rec, dtype=np.int8)
    tags = np.where(tail_ok, auto_tag, b"").tolist()
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

    cat = LineCatalog(
        wavelength=wavelength,
        index_wavelength=wavelength.copy(),
        log_gf=log_gf,
        gf=np.zeros(n_rec),
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
        iso1=iso1.astype(np.int16),
        iso2=iso2.astype(np.int16),
        xj=xj,
        xjp=xjp,
        continuum_label=continuum_label,
        gr_log=gr_log,
        gs_log=gs_log,
        gw_log=gw_log,
    )
    if not np.all(good):
        cat = _merge_whitespace(cat, recs, good)
    # from_records exponentiates the whole catalog's log gf with numpy (atomic.py:275)
    cat.gf = np.power(10.0, cat.log_gf)
    return cat


def _merge_whitespace(cat: LineCatalog, recs, good: np.ndarray) -> LineCatalog:
    """Fixed-width records where good, whitespace-format ones elsewhere, file order."""
    keep, extra = [], {}
    for i in range(len(recs)):
        if good[i]:
            keep.append(i)
        else:
            r = _whitespace_record(recs[i])
            if r is not None:
                keep.append(i)
                extra[i] = r
    keep = np.array(keep, dtype=np.int64)
    out = cat.subset(keep)
    for j, i in enumerate(keep.tolist()):
        r = extra.get(i)
        if r is None:
            continue
        out.wavelength[j] = out.index_wavelength[j] = r["wl"]
        out.log_gf[j] = r["log_gf"]
        out.excitation_energy[j] = r["elo"]
        out.gamma_rad[j], out.gamma_stark[j], out.gamma_vdw[j] = r["gr"], r["gs"], r["gw"]
        out.atomic_number[j] = r["z"]
        out.ion_stages[j] = r["stage"]
        out.line_types[j] = r["lt"]
        out.n_lower[j], out.n_upper[j] = r["nlo"], r["nup"]
        out.code[j] = 0.0
        out.line_size[j] = 0
        out.iso1[j] = out.iso2[j] = 0
        out.xj[j] = out.xjp[j] = 0.0
        out.continuum_label[j] = 0
        out.gr_log[j] = out.gs_log[j] = out.gw_log[j] = 0.0
    return out


def _empty_catalog() -> LineCatalog:
    e = np.zeros(0, dtype=np.float64)
    i16 = np.zeros(0, np.int16)
    return LineCatalog(e, e, e, e, e, e, e, e, np.zeros(0, np.int32), i16, np.zeros(0, np.int8), i16, i16, e,
                       i16, i16, i16, e, e, np.zeros(0, np.uint8), e, e, e)


def as_line_catalog(catalog) -> LineCatalog:
    """This package's LineCatalog for either catalog type.

    A catalog of the reference's own type (io/lines/atomic.py:235-296) carries
    element symbols, not atomic numbers, and keeps CODE and LINESIZE on its
    records; its arrays are taken as they are, Z from the symbol as the engine
    looks it up (engine/opacity.py:634-636)."""
    if isinstance(catalog, LineCatalog):
        return catalog
    from .constants import element_atomic_number

    recs = getattr(catalog, "records", None)
    wl = np.asarray(catalog.wavelength, dtype=np.float64)
    n = wl.size

    def arr(name, dtype):
        v = getattr(catalog, name, None)
        return np.zeros(n, dtype) if v is None else np.asarray(v).astype(dtype)

    def from_records(name, dtype):
        if recs is None:
            return np.zeros(n, dtype)
        return np.array([getattr(r, name, 0) or 0 for r in recs], dtype=dtype)

    idx = getattr(catalog, "index_wavelength", None)
    z = np.array([element_atomic_number(e) or 0 for e in np.asarray(catalog.elements, dtype=object)],
                 dtype=np.int32)
    return LineCatalog(
        wavelength=wl,
        index_wavelength=wl.copy() if idx is None else np.asarray(idx, dtype=np.float64),
        log_gf=arr("log_gf", np.float64),
        gf=arr("gf", np.float64),
        excitation_energy=arr("excitation_energy", np.float64),
        gamma_rad=arr("gamma_rad", np.float64),
        gamma_stark=arr("gamma_stark", np.float64),
        gamma_vdw=arr("gamma_vdw", np.float64),
        atomic_number=z,
        ion_stages=arr("ion_stages", np.int16),
        line_types=arr("line_types", np.int8),
        n_lower=arr("n_lower", np.int16),
        n_upper=arr("n_upper", np.int16),
        code=from_records("code", np.float64),
        line_size=from_records("line_size", np.int16),
        iso1=from_records("iso1", np.int16),
        iso2=from_records("iso2", np.int16),
        xj=from_records("xj", np.float64),
        xjp=from_records("xjp", np.float64),
        continuum_label=np.array([str(getattr(r, "labelp", "") or "").strip().upper().startswith("CONTINUU")
                                  for r in recs] if recs is not None else np.zeros(n), dtype=np.uint8),
        gr_log=from_records("gamma_rad_log", np.float64),
        gs_log=from_records("gamma_stark_log", np.float64),
        gw_log=from_records("gamma_vdw_log", np.float64),
    )


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
    catalog = as_line_catalog(catalog)
    if len(catalog) == 0:
        return catalog
    margin = line_margin(catalog.line_types, catalog.atomic_number, wl_min, catalog.line_size,
                         catalog.code)
    wl = catalog.wavelength
    keep = np.nonzero((wl >= wl_min - margin) & (wl <= wl_max + margin))[0]
    return catalog.subset(keep)


def write_gfall(path: Path, wl_nm: np.ndarray, log_gf: np.ndarray, code: np.ndarray,
                e_low_cm: np.ndarray, gr_log: np.ndarray, gs_log: np.ndarray,
                gw_log: np.ndarray, n_lower: Optional[np.ndarray] = None,
                n_upper: Optional[np.ndarray] = None) -> None:
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
            f"{gr_log[i]:6.2f}{gs_log[i]:6.2f}{gw_log[i]:6.2f}{'SYN ':4s}"
            f"{0 if n_lower is None else int(n_lower[i]):2d}{0 if n_upper is None else int(n_upper[i]):2d}"
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
    """Host prep of the per-line arrays (line_opacity.py:740-784, 1030-1059).

    catalog is this package's LineCatalog or the reference's (as_line_catalog)."""
    catalog = as_line_catalog(catalog)
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


@dataclass
class HydrogenLines:
    """The lines engine/opacity.py:1776-1782 sends to the HPROF4 path (line
    type -1/-2, ion stage 1), in catalog order, as synthe_hydrogen_lines reads
    them: the record's wavelength, the centre index of the index wavelength
    (engine/opacity.py:550-592, -1 / W off the grid), linear gf, lower level
    energy and NBLO/NBUP (opacity.py:1787-1796 defaults when absent)."""

    index: np.ndarray
    wavelength: np.ndarray
    center: np.ndarray
    gf: np.ndarray
    elo: np.ndarray
    n_lower: np.ndarray
    n_upper: np.ndarray

    def __len__(self) -> int:
        return int(self.index.size)


def hydrogen_lines(catalog, grid: np.ndarray) -> HydrogenLines:
    catalog = as_line_catalog(catalog)
    lt = np.asarray(catalog.line_types, np.int64)
    idx = np.nonzero(((lt == -1) | (lt == -2)) & (np.asarray(catalog.ion_stages, np.int64) == 1))[0]
    if idx.size and grid.size:
        center = nearest_grid_indices(grid, np.asarray(catalog.index_wavelength, np.float64)[idx])
    else:
        center = np.zeros(idx.size, np.int64)
    if catalog.n_lower is not None:
        nlo = np.asarray(catalog.n_lower, np.int64)[idx]
    else:
        nlo = np.ones(idx.size, np.int64)
    if catalog.n_upper is not None:
        nup = np.asarray(catalog.n_upper, np.int64)[idx]
    else:
        nup = nlo + 1
    c = lambda a, t: np.ascontiguousarray(a, t)  # noqa: E731
    return HydrogenLines(idx, c(np.asarray(catalog.wavelength)[idx], np.float64), c(center, np.int64),
                         c(np.asarray(catalog.gf)[idx], np.float64),
                         c(np.asarray(catalog.excitation_energy)[idx], np.float64), c(nlo, np.int32),
                         c(nup, np.int32))


# ------------------------------------------------------------- fort.19 lines
# Element/ion codes with a WCON table column, 1-based (io/lines/fort19.py:19-22)
_CODEX = np.array([1.0, 2.0, 2.01, 6.0, 6.01, 12.0, 12.01, 13.0, 13.01, 14.0, 14.01, 20.0, 20.01, 8.0, 11.0,
                   5.0, 19.0], dtype=np.float64)
# fort.19 line kinds the ASYNTH add handles (engine/opacity.py:1374-1381)
F19_NORMAL, F19_AUTOIONIZING, F19_CONTINUUM = 0, 1, 2


@dataclass
class Fort19Records:
    """The fort.19 wing records rgfall writes for a window (io/lines/fort19.py:73-90),
    in catalog order, with the catalog index each was built from.  The float
    fields are float32, as the record stores them."""

    source: np.ndarray  # int64 catalog index
    wavelength_vacuum: np.ndarray  # float64, nm
    energy_lower: np.ndarray  # float32
    oscillator_strength: np.ndarray  # float32: CGF for TYPE <= 3 (not 1), else gf
    n_lower: np.ndarray
    n_upper: np.ndarray
    ion_index: np.ndarray  # NELION
    line_type: np.ndarray
    continuum_index: np.ndarray  # NCON
    element_index: np.ndarray  # NELIONX
    gamma_rad: np.ndarray  # float32
    gamma_stark: np.ndarray
    gamma_vdw: np.ndarray
    nbuff: np.ndarray
    limb: np.ndarray

    def __len__(self) -> int:
        return int(self.wavelength_vacuum.size)


def _opt(catalog: LineCatalog, name: str, dtype) -> np.ndarray:
    v = getattr(catalog, name, None)
    return np.zeros(len(catalog), dtype) if v is None else np.asarray(v).astype(dtype)


def fort19_records(catalog, wlbeg: float, wlend: float, resolution: float) -> Fort19Records:
    """rgfall's fort.19 selection and fields (io/lines/fort19.py:233-420).

    A record goes to fort.19 when its TYPE is 1 (autoionising) or above 3
    (merged continuum: LABELP "CONTINUU", TYPE = the last member NLAST, gf
    times 2J + 1), or when it carries NBLO/NBUP; coronal lines (TYPE 2) are
    dropped; the DELLIM margin is the window filter's.  TYPE 1 and > 3 keep
    their damping on the file's unnormalised scale (a positive log GS meaning
    a negative Shore A); the others keep the catalog's normalised constants
    and store CGF in place of gf.  The float fields are rounded to float32."""
    catalog = as_line_catalog(catalog)
    n = len(catalog)
    ratiolg = math.log(1.0 + 1.0 / resolution)
    ixwlbeg = math.floor(math.log(wlbeg) / ratiolg)
    if math.exp(ixwlbeg * ratiolg) < wlbeg:
        ixwlbeg += 1
    delfactor = 1.0 if wlbeg <= 500.0 else wlbeg / 500.0
    wl = np.asarray(catalog.wavelength, np.float64)
    code_raw = _opt(catalog, "code", np.float64)
    size = np.maximum(_opt(catalog, "line_size", np.int64), 0)
    lim = np.minimum(8 - size, 7)
    lim = np.where(np.abs(np.where(code_raw > 0.0, code_raw, 0.0) - 1.0) < 1.0e-6, 1, lim)
    margin = _DELLIM[lim - 1] * delfactor
    nblo = np.abs(_opt(catalog, "n_lower", np.int64))
    nbup = np.abs(_opt(catalog, "n_upper", np.int64))
    lt = _opt(catalog, "line_types", np.int64)
    cont = _opt(catalog, "continuum_label", np.uint8) != 0
    xj, xjp = _opt(catalog, "xj", np.float64), _opt(catalog, "xjp", np.float64)
    ltype = np.where(cont, np.where(xjp > 0.0, xjp.astype(np.int64), nbup), lt)
    cand = (wl >= wlbeg - margin) & (wl <= wlend + margin) & (ltype != 2)
    cand &= (ltype == 1) | (ltype > 3) | (nblo + nbup != 0)
    idx = np.nonzero(cand)[0]
    gr_log, gs_log, gw_log = (_opt(catalog, k, np.float64) for k in ("gr_log", "gs_log", "gw_log"))
    z = _opt(catalog, "atomic_number", np.int64)
    stage = _opt(catalog, "ion_stages", np.int64)
    iso1, iso2 = _opt(catalog, "iso1", np.int64), _opt(catalog, "iso2", np.int64)
    cols = {k: [] for k in ("gf", "nelion", "nelionx", "ncon", "gr", "gs", "gw", "nbuff")}
    for i in idx.tolist():  # the reference's scalar arithmetic, record by record
        wlvac = float(wl[i])
        t = int(ltype[i])
        gf = 10.0 ** float(catalog.log_gf[i])
        if cont[i]:
            gf *= 2.0 * float(xj[i]) + 1.0
        code = float(code_raw[i]) if code_raw[i] > 0.0 else 0.0
        if code <= 0.0 and z[i] > 0:  # CODE from the element and stage
            code = int(z[i]) + (int(stage[i]) - 1) * 0.01
        nelem = int(code + 1.0e-6) if code > 0.0 else 0
        icharge = int((code - nelem) * 100.0 + 0.1) if code > 0.0 else 0
        nelion = nelem * 6 - 6 + icharge + 1
        if 19 < nelem < 29 and icharge > 5:
            nelion = 6 * (nelem + icharge * 10 - 30) - 1
        nelionx = 0
        if code > 0.0:
            hit = np.nonzero(np.isclose(_CODEX, code, rtol=0.0, atol=1e-3))[0]
            nelionx = int(hit[0]) + 1 if hit.size else 0
        freq = C_LIGHT_NM / max(wlvac, 1.0e-30)
        denom = 12.5664 * freq
        if t == 1 or t > 3:
            gr = 10.0 ** gr_log[i] if gr_log[i] != 0.0 else catalog.gamma_rad[i] * denom
            if gs_log[i] > 0.0:
                gs = -10.0 ** (-gs_log[i])
            elif gs_log[i] < 0.0:
                gs = 10.0 ** gs_log[i]
            else:
                gs = catalog.gamma_stark[i] * denom
            gw = 10.0 ** gw_log[i] if gw_log[i] != 0.0 else catalog.gamma_vdw[i] * denom
        else:
            gr, gs, gw = catalog.gamma_rad[i], catalog.gamma_stark[i], catalog.gamma_vdw[i]
            gf = CGF_CONSTANT * gf / freq
        cols["gf"].append(gf)
        cols["nelion"].append(nelion)
        cols["nelionx"].append(nelionx)
        cols["ncon"].append(int(iso2[i]) if iso1[i] == 0 and iso2[i] > 0 else 0)
        cols["gr"].append(float(gr))
        cols["gs"].append(float(gs))
        cols["gw"].append(float(gw))
        cols["nbuff"].append(int(math.log(wlvac) / ratiolg + 0.5) - int(ixwlbeg) + 1)
    f4 = lambda k: np.asarray(cols[k], np.float32)  # noqa: E731
    return Fort19Records(
        source=idx.astype(np.int64), wavelength_vacuum=wl[idx].copy(),
        energy_lower=np.asarray(catalog.excitation_energy, np.float64)[idx].astype(np.float32),
        oscillator_strength=f4("gf"), n_lower=nblo[idx].astype(np.int16), n_upper=nbup[idx].astype(np.int16),
        ion_index=np.asarray(cols["nelion"], np.int16), line_type=ltype[idx].astype(np.int16),
        continuum_index=np.asarray(cols["ncon"], np.int16), element_index=np.asarray(cols["nelionx"], np.int16),
        gamma_rad=f4("gr"), gamma_stark=f4("gs"), gamma_vdw=f4("gw"),
        nbuff=np.asarray(cols["nbuff"], np.int32), limb=lim[idx].astype(np.int32))


def match_catalog_to_fort19(catalog_wavelength: np.ndarray, meta_wavelength: np.ndarray,
                            tolerance: float = 1e-3) -> dict:
    """catalog index -> fort.19 index (engine/opacity.py:595-620): each record
    takes the nearer of its two sorted neighbours within the tolerance, the
    first record to claim a catalog line keeps it."""
    mapping = {}
    cw = np.asarray(catalog_wavelength, np.float64)
    mw = np.asarray(meta_wavelength, np.float64)
    if cw.size == 0 or mw.size == 0:
        return mapping
    order = np.argsort(cw)
    sw = cw[order]
    for m, pos in enumerate(np.searchsorted(sw, mw).tolist()):
        best, best_d = None, tolerance
        for cand in (pos, pos - 1):
            if 0 <= cand < sw.size:
                dlt = abs(sw[cand] - mw[m])
                if dlt < best_d:
                    best_d, best = dlt, cand
        if best is not None:
            ci = int(order[best])
            if ci not in mapping:
                mapping[ci] = m
    return mapping


# WCON table values by (NCON, NELIONX), cm^-1 (physics/tables.py:163-204; zero rows omitted)
def _contx() -> np.ndarray:
    with np.load(Path(__file__).resolve().parent / "data" / "kurucz_tables.npz", allow_pickle=False) as z:
        return np.asarray(z["contx"], np.float64)


@dataclass
class Fort19Lines:
    """The fort.19 records the ASYNTH add takes (engine/opacity.py:1364-1397),
    in fort.19 order, as synthe_fort19_lines reads them: kind (F19_*), the
    record's wavelength and centre index (engine/opacity.py:550-592), its
    float32 gf and damping as doubles, the matched catalog line's wavelength
    (the Doppler width's scale), lower level energy, Z - 1 and stage, the
    record's NELION and NLAST, and CONTX(NCON, NELIONX) (0 where no WCON)."""

    kind: np.ndarray
    wavelength: np.ndarray
    center: np.ndarray
    gf: np.ndarray
    gamma_rad: np.ndarray
    gamma_stark: np.ndarray
    gamma_vdw: np.ndarray
    cat_wavelength: np.ndarray
    elo: np.ndarray
    elem: np.ndarray
    ion: np.ndarray
    nelion: np.ndarray
    nlast: np.ndarray
    cont_val: np.ndarray
    cat_index: np.ndarray  # the matched catalog line, for per-line population quantities

    def __len__(self) -> int:
        return int(self.kind.size)


def fort19_lines(catalog, records, grid: np.ndarray, n_ion: int, n_elem: int,
                 catalog_to_fort19: Optional[dict] = None, contx: Optional[np.ndarray] = None) -> Fort19Lines:
    """Select and pack the records (engine/opacity.py:1364-1397): wing types
    normal, autoionising and merged continuum, matched to a catalog line of a
    known element and positive stage, centre not past the grid's end.  A
    matched line whose element or stage lies outside the bundle's population
    table raises IndexError, where the reference's table lookup fails.

    records: Fort19Records or the reference's Fort19Data (the same field
    names); catalog_to_fort19: the engine's match (computed here when None);
    contx: the WCON table (the reference's metal_wing_tables().contx by default)."""
    catalog = as_line_catalog(catalog)
    grid = np.asarray(grid, np.float64)
    rwl = np.asarray(records.wavelength_vacuum, np.float64)
    if catalog_to_fort19 is None:
        catalog_to_fort19 = match_catalog_to_fort19(catalog.wavelength, rwl)
    to_cat = {v: k for k, v in catalog_to_fort19.items()}
    centers = nearest_grid_indices(grid, rwl) if rwl.size else np.zeros(0, np.int64)
    contx = _contx() if contx is None else np.asarray(contx, np.float64)
    out = {k: [] for k in Fort19Lines.__dataclass_fields__}
    for f in range(rwl.size):
        t = int(records.line_type[f])
        kind = F19_NORMAL if t == 0 else F19_AUTOIONIZING if t == 1 else F19_CONTINUUM if t > 3 else -1
        ci = to_cat.get(f)
        if kind < 0 or ci is None:
            continue
        zc, st = int(catalog.atomic_number[ci]), int(catalog.ion_stages[ci])
        if zc <= 0 or st <= 0:
            continue
        if st - 1 >= n_ion or zc - 1 >= n_elem:
            raise IndexError(f"fort.19 record {f}: element {zc} stage {st} outside the population table "
                             f"({n_ion} stages, {n_elem} elements)")
        c = int(centers[f])
        if c >= grid.size:
            continue
        ncon, nx = int(records.continuum_index[f]), int(records.element_index[f])
        cv = 0.0
        if kind == F19_NORMAL and ncon > 0 and nx > 0 and nx <= contx.shape[1] and ncon <= contx.shape[0]:
            cv = float(contx[ncon - 1, nx - 1])
        vals = dict(kind=kind, wavelength=float(rwl[f]), center=c,
                    gf=float(records.oscillator_strength[f]), gamma_rad=float(records.gamma_rad[f]),
                    gamma_stark=float(records.gamma_stark[f]), gamma_vdw=float(records.gamma_vdw[f]),
                    cat_wavelength=float(catalog.wavelength[ci]), elo=float(catalog.excitation_energy[ci]),
                    elem=zc - 1, ion=st, nelion=int(records.ion_index[f]), nlast=t, cont_val=cv, cat_index=ci)
        for k, v in vals.items():
            out[k].append(v)
    i32, f8 = np.int32, np.float64
    types = dict(kind=i32, center=np.int64, elem=i32, ion=i32, nelion=i32, nlast=i32, cat_index=np.int64)
    return Fort19Lines(**{k: np.ascontiguousarray(np.asarray(v, types.get(k, f8))) for k, v in out.items()})


def helium_wing_lines(catalog, grid: np.ndarray, records=None, n_ion: int = 6,
                      contx: Optional[np.ndarray] = None) -> Fort19Lines:
    """The lines of the engine's helium pass (engine/opacity.py:2534-2546 and
    2686-3066), packed for synthe_fort19_lines as kind 3, in catalog order:
    TYPE -3, -4 or -6 (the matched fort.19 record's TYPE where there is one),
    centre (of the catalog wavelength) on the grid, ion stage within the
    population table.  gf holds CGF = 0.026538/1.77245 gf / nu, nelion the
    ion stage, nlast the TYPE, cont_val CONTX of the matched record's NCON and
    NELIONX; records: fort19_records of the window (matched here)."""
    catalog = as_line_catalog(catalog)
    grid = np.asarray(grid, np.float64)
    lt = np.asarray(catalog.line_types, np.int64).copy()
    ncon = np.zeros(lt.size, np.int64)
    nelionx = np.zeros(lt.size, np.int64)
    if records is not None and len(records):
        for ci, f in match_catalog_to_fort19(catalog.wavelength, records.wavelength_vacuum).items():
            lt[ci] = int(records.line_type[f])
            ncon[ci] = int(records.continuum_index[f])
            nelionx[ci] = int(records.element_index[f])
    wl = np.asarray(catalog.wavelength, np.float64)
    z = np.asarray(catalog.atomic_number, np.int64)
    st = np.asarray(catalog.ion_stages, np.int64)
    centers = nearest_grid_indices(grid, wl) if wl.size else np.zeros(0, np.int64)
    pick = np.isin(lt, (-3, -4, -6)) & (centers >= 0) & (centers < grid.size) & ~((z == 1) & (st == 1))
    pick &= (st >= 1) & (st <= n_ion) & (z > 0)
    idx = np.nonzero(pick)[0]
    contx = _contx() if contx is None else np.asarray(contx, np.float64)
    cv = np.zeros(idx.size)
    for j, i in enumerate(idx.tolist()):
        if ncon[i] > 0 and nelionx[i] > 0 and nelionx[i] <= contx.shape[1] and ncon[i] <= contx.shape[0]:
            cv[j] = contx[ncon[i] - 1, nelionx[i] - 1]
    w = wl[idx]
    cgf = CGF_CONSTANT * np.asarray(catalog.gf, np.float64)[idx] / (C_LIGHT_NM / w)
    c = np.ascontiguousarray
    return Fort19Lines(kind=np.full(idx.size, 3, np.int32), wavelength=c(w), center=c(centers[idx], np.int64),
                       gf=c(cgf), gamma_rad=c(np.asarray(catalog.gamma_rad, np.float64)[idx]),
                       gamma_stark=c(np.asarray(catalog.gamma_stark, np.float64)[idx]),
                       gamma_vdw=c(np.asarray(catalog.gamma_vdw, np.float64)[idx]), cat_wavelength=c(w),
                       elo=c(np.asarray(catalog.excitation_energy, np.float64)[idx]),
                       elem=c(z[idx] - 1, np.int32), ion=c(st[idx], np.int32), nelion=c(st[idx], np.int32),
                       nlast=c(lt[idx], np.int32), cont_val=c(cv), cat_index=c(idx, np.int64))
