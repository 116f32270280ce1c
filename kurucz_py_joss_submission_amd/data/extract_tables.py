"""Write ``kurucz_tables.npz``: the numeric tables the synthe hot path needs.

The tables are Kurucz ATLAS/SYNTHE data, not code: the JOSH quadrature
weights and Lambda-operator matrices (``synthe_py/physics/josh_tables.py``),
the Voigt H1 node table (``synthe_py/physics/tables.py:249-272``), the
POTION ionisation potentials in cm^-1 that the GFALL damping defaults use
(``synthe_py/data/fortran_data.npz``, read by ``io/lines/atomic.py:21-37``),
the HPROF4 hydrogen-profile tables (``physics/profiles/hydrogen.py:47-1406``:
Stark PROPBM corrections, C/D coefficients, fine-structure components,
quasi-static cutoffs, radiative damping sums, XKNMTB, Y1WTM), the hydrogen
series limits CONTH (``physics/tables.py:152``), the WCON merge table CONTX
(``physics/tables.py:163-204``) and the EHYD level energies
(``engine/opacity.py:49-61``).
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


def _array_literal(node: ast.AST) -> np.ndarray | None:
    """np.array(<literal>) or np.array(<literal>).reshape(a, b, order="F")."""
    if not isinstance(node, ast.Call):
        return None
    f = node.func
    if isinstance(f, ast.Attribute) and f.attr == "reshape":
        base = _array_literal(f.value)
        if base is None:
            return None
        shape = tuple(ast.literal_eval(a) for a in node.args)
        order = "C"
        for kw in node.keywords:
            if kw.arg == "order":
                order = ast.literal_eval(kw.value)
        return base.reshape(shape, order=order)
    if isinstance(f, ast.Attribute) and f.attr == "array" and node.args:
        try:
            return np.asarray(ast.literal_eval(node.args[0]), dtype=np.float64)
        except ValueError:
            return None
    return None


def _hydrogen_tables(path: Path) -> dict[str, np.ndarray]:
    """The literal tables inside ``hydrogen_tables()``; the three stacked
    ones (PROPBM, C, D) are stacked here as that function stacks them."""
    tree = ast.parse(path.read_text())
    fn = next(n for n in ast.walk(tree) if isinstance(n, ast.FunctionDef) and n.name == "hydrogen_tables")
    found: dict[str, object] = {}
    for node in fn.body:
        if not isinstance(node, ast.Assign) or len(node.targets) != 1 or not isinstance(node.targets[0], ast.Name):
            continue
        name = node.targets[0].id
        if isinstance(node.value, ast.List):
            found[name] = [_array_literal(e) for e in node.value.elts]
        else:
            arr = _array_literal(node.value)
            if arr is not None:
                found[name] = arr
    out = {k: v for k, v in found.items() if isinstance(v, np.ndarray)}
    out["propbm"] = np.stack([a.reshape(5, 15, order="F") for a in found["prob_arrays"]], axis=0)
    out["c"] = np.stack(found["c_cols"], axis=1)
    out["d"] = np.stack(found["d_cols"], axis=1)
    return out


def main(ref_root: str = "/root/reference") -> None:
    root = Path(ref_root) / "synthe_py" / "physics"
    josh = _literal_assignments(root / "josh_tables.py")
    voigt = _literal_assignments(root / "tables.py")  # tabvi / tabh1 inside _voigt_reference_tables
    with np.load(Path(ref_root) / "synthe_py" / "data" / "fortran_data.npz", allow_pickle=False) as z:
        potion = np.asarray(z["potion"], dtype=np.float64)
    hyd = _hydrogen_tables(root / "profiles" / "hydrogen.py")
    engine = _literal_assignments(Path(ref_root) / "synthe_py" / "engine" / "opacity.py")
    tables = {
        "xtau": josh["XTAU_GRID"],
        "ch": josh["CH_WEIGHTS"],
        "ck": josh["CK_WEIGHTS"],
        "coefj": josh["COEFJ_MATRIX"],
        "coefh": josh["COEFH_MATRIX"],
        "tabvi": voigt["tabvi"],
        "tabh1": voigt["tabh1"],
        "potion": potion,
        "conth": voigt["_METAL_CONTH"],
        "contx": voigt["_METAL_CONTX"],
        "ehyd_cm": engine["_EHYD_CM"],
    }
    for k in ("propbm", "c", "d", "pp", "beta_grid", "stalph", "stwtal", "istal", "lnghal", "stcomp", "stcpwt",
              "lncomp", "cutoff_h2_plus", "cutoff_h2", "asum_lyman", "asum", "y1wtm", "xknmtb"):
        tables["h_" + k] = hyd[k]
    assert tables["h_propbm"].shape == (7, 5, 15) and tables["h_c"].shape == (5, 7)
    assert tables["h_stcomp"].shape == (5, 4) and tables["h_xknmtb"].shape == (4, 3)
    assert tables["ehyd_cm"].shape == (8,)
    assert tables["contx"].shape == (26, 17)
    assert tables["xtau"].shape == (51,) and tables["coefj"].shape == (51, 51)
    assert tables["tabvi"].shape == (81,) and tables["tabh1"].shape == (81,)
    assert tables["potion"].shape == (999,)
    np.savez(OUT, **tables)
    print(f"wrote {OUT}: " + ", ".join(f"{k}{v.shape}" for k, v in tables.items()))


if __name__ == "__main__":
    main(*sys.argv[1:])
