"""Hand-written GFALL records for the fort.19 tests (fixed-width columns of
io/lines/atomic.py:325-345): every field the fort.19 routing reads --
LABELP, the J values, NBLO/NBUP, ISO1/ISO2, the LINESIZE digit and the
AUT/COR/PRD tag -- can be set."""

from __future__ import annotations


def record(wl_nm: float, log_gf: float, code: float, e_low: float, gr: float = 8.0, gs: float = -5.0,
           gw: float = -7.5, xj: float = 0.5, xjp: float = 1.5, labelp: str = "UPLEV", n_lower: int = 0,
           n_upper: int = 0, iso1: int = 0, iso2: int = 0, tag: str = "", line_size: str = " ") -> str:
    """One 160-column record with EP = E + 1e7 / wl (so the parsed wavelength is wl)."""
    ep = e_low + 1.0e7 / wl_nm
    rec = (f"{wl_nm:11.4f}{log_gf:7.3f}{code:6.2f}{e_low:12.3f}{xj:5.1f} {'LOWLEV':8s}{'  ':2s}"
           f"{ep:12.3f}{xjp:5.1f} {labelp:8s}{'  ':2s}{gr:6.2f}{gs:6.2f}{gw:6.2f}{'SYN ':4s}"
           f"{n_lower:2d}{n_upper:2d}{iso1:3d}{0.0:6.3f}{iso2:3d}{0.0:6.3f}{'':10s}"
           f"{'':6s}{line_size:1s}{tag:3s}{'':10s}{0:6d}")
    assert len(rec) == 160, len(rec)
    return rec
