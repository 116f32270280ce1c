"""Switch a synthe_py installation's hot path to this package's drop-ins.

The reference engine reaches the hot path in two ways
(synthe_py/engine/opacity.py:23-35, 2236, 2320, 2352, 3788):

* through module attributes, ``continuum.build_depth_continuum(...)``,
  ``line_opacity.compute_transp(...)`` and
  ``line_opacity.compute_asynth_from_transp(...)``, so those attributes of the
  physics modules are replaced;
* through names bound at import, ``from .radiative import solve_lte_spectrum``
  in engine/opacity.py and ``from synthe_py.physics.josh_solver import
  solve_josh_flux`` in engine/radiative.py (radiative.py:12, 192, 208), so the
  name is replaced in every module that holds it.

``install()`` does both and returns the replaced originals; ``uninstall()``
puts them back.
"""

from __future__ import annotations

import importlib
from typing import Dict, Optional, Tuple

# (module, attribute, drop-in name)
BINDINGS: Tuple[Tuple[str, str, str], ...] = (
    ("physics.line_opacity", "compute_transp", "compute_transp"),
    ("physics.line_opacity", "compute_asynth_from_transp", "compute_asynth_from_transp"),
    ("physics.continuum", "build_depth_continuum", "build_depth_continuum"),
    ("physics.josh_solver", "solve_josh_flux", "solve_josh_flux"),
    ("engine.radiative", "solve_josh_flux", "solve_josh_flux"),
    ("engine.radiative", "solve_lte_spectrum", "solve_lte_spectrum"),
    ("engine.radiative", "solve_lte_frequency", "solve_lte_frequency"),
    ("engine.opacity", "solve_lte_spectrum", "solve_lte_spectrum"),
    ("physics.bfudge", "compute_bfudge_and_slinec", "compute_bfudge_and_slinec"),
    ("engine.opacity", "_compute_hydrogen_line_opacity", "compute_hydrogen_line_opacity"),
    ("engine.opacity", "_add_fort19_asynth", "add_fort19_asynth"),
)


def install(package: str = "synthe_py", impl=None) -> Dict[Tuple[str, str], object]:
    """Point the reference package's hot-path functions at the drop-ins.

    impl: the module holding the drop-ins (default: this package's physics
    module, whose entry points raise NativeUnavailable without a GPU).
    Modules of the package that are not importable are skipped."""
    if impl is None:
        from . import physics as impl
    saved: Dict[Tuple[str, str], object] = {}
    for mod_name, attr, name in BINDINGS:
        try:
            mod = importlib.import_module(f"{package}.{mod_name}")
        except ImportError:
            continue
        if not hasattr(mod, attr):
            continue
        saved[(f"{package}.{mod_name}", attr)] = getattr(mod, attr)
        setattr(mod, attr, getattr(impl, name))
    return saved


def uninstall(saved: Dict[Tuple[str, str], object]) -> None:
    for (mod_name, attr), fn in saved.items():
        setattr(importlib.import_module(mod_name), attr, fn)


def install_from_env(package: str = "synthe_py") -> Optional[Dict[Tuple[str, str], object]]:
    """install() when SYNTHE_BACKEND=mi355x, for a one-line hook in the reference."""
    import os

    if os.environ.get("SYNTHE_BACKEND") != "mi355x":
        return None
    return install(package)
