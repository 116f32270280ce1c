"""Generate _tables_gen.h (C arrays) from data/kurucz_tables.npz for the HIP library.

Values are printed with 17 significant digits, which round-trips IEEE doubles
exactly, so the library's tables are bit-identical to the archive.
"""

from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
NPZ = HERE.parent / "data" / "kurucz_tables.npz"
OUT = HERE / "_tables_gen.h"


def _arr(name: str, a: np.ndarray, device: bool = False) -> str:
    vals = ",\n  ".join(", ".join(f"{float(v):.17g}" for v in a.ravel()[i:i + 4])
                        for i in range(0, a.size, 4))
    q = "static __device__ const double" if device else "static const double"
    return f"{q} {name}[{a.size}] = {{\n  {vals}\n}};\n"


def main(out: Path = OUT) -> Path:
    t = np.load(NPZ)
    parts = ["// Generated by gen_tables.py from data/kurucz_tables.npz; do not edit.\n",
             "#pragma once\n"]
    for key, name in (("xtau", "kXtau"), ("ch", "kCh"), ("coefj", "kCoefj"), ("tabvi", "kTabvi"),
                      ("tabh1", "kTabh1")):
        parts.append(_arr(name, np.asarray(t[key], dtype=np.float64)))
    # hydrogen (HPROF4) tables, row-major as the archive holds them: host
    # copies for the per-line constants, device copies for the per-pixel profile
    for key in ("h_stalph", "h_stwtal", "h_istal", "h_lnghal", "h_stcomp", "h_stcpwt", "h_lncomp", "h_asum_lyman",
                "h_asum", "h_y1wtm", "h_xknmtb", "conth", "ehyd_cm"):
        parts.append(_arr("kH_" + key.removeprefix("h_"), np.asarray(t[key], dtype=np.float64)))
    for key in ("h_propbm", "h_c", "h_d", "h_pp", "h_beta_grid", "h_cutoff_h2", "h_cutoff_h2_plus"):
        parts.append(_arr("kD_" + key.removeprefix("h_"), np.asarray(t[key], dtype=np.float64), device=True))
    text = "\n".join(parts)
    if not out.exists() or out.read_text() != text:
        out.write_text(text)
    return out


if __name__ == "__main__":
    main(Path(sys.argv[1]) if len(sys.argv) > 1 else OUT)
