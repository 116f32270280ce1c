"""The spectrv control card (synthe_py/io/spectrv.py): the first line of a
spectrv input deck holds eight floats, RHOXJ R1 R101 PH1 PC1 PSI1 PRDDOP
PRDPOW, of which the synthesis uses RHOXJ (line scattering) and the BFUDGE
exponents PH1, PC1, PSI1."""

from __future__ import annotations

from dataclasses import dataclass
from pathlib import Path


@dataclass(frozen=True)
class SpectrvParams:
    """io/spectrv.py:8-17."""

    rhoxj: float
    ph1: float
    pc1: float
    psi1: float
    prddop: float
    prdpow: float


LTE = SpectrvParams(rhoxj=0.0, ph1=0.0, pc1=0.0, psi1=0.0, prddop=0.0, prdpow=0.0)


def load(path: Path) -> SpectrvParams:
    """Parse a spectrv input deck (io/spectrv.py:20-44): the first line's eight
    floats, separated by blanks or commas; ValueError on an empty deck or a
    short line."""
    lines = Path(path).read_text(encoding="ascii").splitlines()
    if not lines or not lines[0].strip():
        raise ValueError(f"spectrv input {path} is empty or has no parameter line")
    parts = lines[0].replace(",", " ").split()
    if len(parts) < 8:
        raise ValueError(f"spectrv input {path} expected eight floats on first line, found {parts}")
    rhoxj, _r1, _r101, ph1, pc1, psi1, prddop, prdpow = (float(x) for x in parts[:8])
    return SpectrvParams(rhoxj=rhoxj, ph1=ph1, pc1=pc1, psi1=psi1, prddop=prddop, prdpow=prdpow)
