"""Write ``kurucz_tables.npz``: the numeric tables the synthe hot path needs.

The tables are Kurucz ATLAS/SYNTHE data, not code: the JOSH quadrature
weights and Lambda-operator matrices (``synthe_py/physics/josh_tables.py``),
the Voigt H1 node table (``synthe_py/physics/tables.py:249-272``) and the
POTION ionisation potentials in cm^-1 that the GFALL damping defaults use
(``synthe_py/data/fortran_data.npz``, read by ``io/lines/atomic.py:21-37``).
The source tables are read as text with ``ast`` and the archive with
``numpy.load(allow_pickle=False)`` (nothing from the reference is imported
or executed) and stored as float64 arrays, so the package and the GPU
library carry the identical numbers without the reference present.

Usage::

    python kurucz_py_joss_submission_amd/data/extract_tables.py [REFERENCE_ROOT]

REFERENCE_ROOT defaults to /root/reference.  The archive is committed, so this
only needs re-running if the reference tables change.
"""

from __future__ import annotations

import ast
import sys
from pathlib import Path

import numpy as np

OUT = Path(__file__).resolve().parent / "kurucz_tables.npz"


def _literal_assignments(path: Path) -> dict[str, np.ndarray]:
    """Return every ``NAME = np.array([...], ...)`` literal in ``path``."""
    tree = ast.parse(path.read_text())
    found: dict[str, np.ndarray] = {}
    for node in ast.walk(tree):
        if not isinstance(node, ast.Assign) or len(node.targets) != 1:
            continue
        target = node.targets[0]
        if not isinstance(target, ast.Name):
            continue
        call = node.value
        if isinstance(call, ast.Call) and call.args:
            try:
                found[target.id] = np.asarray(ast.literal_eval(call.args[0]), dtype=np.float64)
            except ValueError:
                continue
    return found


def main(ref_root: str = "/root/reference") -> None:
    root = Path(ref_root) / "synthe_py" / "physics"
    josh = _literal_assignments(root / "josh_tables.py")
    voigt = _literal_assignments(root / "tables.py")  # tabvi / tabh1 inside _voigt_reference_tables
    with np.load(Path(ref_root) / "synthe_py" / "data" / "fortran_data.npz", allow_pickle=False) as z:
        potion = np.asarray(z["potion"], dtype=np.float64)
    tables = {
        "xtau": josh["XTAU_GRID"],
        "ch": josh["CH_WEIGHTS"],
        "ck": josh["CK_WEIGHTS"],
        "coefj": josh["COEFJ_MATRIX"],
        "coefh": josh["COEFH_MATRIX"],
        "tabvi": voigt["tabvi"],
        "tabh1": voigt["tabh1"],
        "potion": potion,
    }
    assert tables["xtau"].shape == (51,) and tables["coefj"].shape == (51, 51)
    assert tables["tabvi"].shape == (81,) and tables["tabh1"].shape == (81,)
    assert tables["potion"].shape == (999,)
    np.savez(OUT, **tables)
    print(f"wrote {OUT}: " + ", ".join(f"{k}{v.shape}" for k, v in tables.items()))


if __name__ == "__main__":
    main(*sys.argv[1:])
