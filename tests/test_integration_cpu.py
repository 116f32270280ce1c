"""integration.install() on stand-in modules laid out like the reference
(synthe_py/engine/opacity.py:23-35 imports the physics modules and binds
solve_lte_spectrum by name; engine/radiative.py:12 binds solve_josh_flux):
every hot-path call the engine makes reaches the drop-in after install and
the original after uninstall.  No GPU: the drop-ins are recording stand-ins."""

from __future__ import annotations

import sys
import types

import pytest

PKG = "synthe_ref_layout"


def _module(name, **attrs):
    m = types.ModuleType(name)
    m.__dict__.update(attrs)
    sys.modules[name] = m
    return m


@pytest.fixture
def ref_layout():
    calls = []

    def orig(tag):
        def f(*a, **k):
            calls.append(("ref", tag))
            return tag
        return f

    pkg = _module(PKG)
    pkg.__path__ = []
    phys = _module(f"{PKG}.physics")
    phys.__path__ = []
    eng = _module(f"{PKG}.engine")
    eng.__path__ = []
    lo = _module(f"{PKG}.physics.line_opacity", compute_transp=orig("transp"),
                 compute_asynth_from_transp=orig("asynth"))
    co = _module(f"{PKG}.physics.continuum", build_depth_continuum=orig("continuum"))
    js = _module(f"{PKG}.physics.josh_solver", solve_josh_flux=orig("josh"))
    bf = _module(f"{PKG}.physics.bfudge", compute_bfudge_and_slinec=orig("bfudge"))
    # engine/radiative.py:12: `from synthe_py.physics.josh_solver import solve_josh_flux`
    rad = _module(f"{PKG}.engine.radiative", solve_josh_flux=js.solve_josh_flux,
                  solve_lte_spectrum=orig("lte_spectrum"), solve_lte_frequency=orig("lte_frequency"))
    # engine/opacity.py:23-35: module imports plus `from .radiative import solve_lte_spectrum`
    # and its own _compute_hydrogen_line_opacity (:1706) and _add_fort19_asynth (:1335),
    # called by name at :2467 and :2375
    op = _module(f"{PKG}.engine.opacity", line_opacity=lo, continuum=co, bfudge=bf,
                 solve_lte_spectrum=rad.solve_lte_spectrum, _compute_hydrogen_line_opacity=orig("hydrogen"),
                 _add_fort19_asynth=orig("fort19"))

    def run_synthesis():  # the engine's call sites (2236, 2320, 2352, 2375, 2467, 3788)
        g = op.__dict__
        return [g["continuum"].build_depth_continuum(), g["line_opacity"].compute_transp(),
                g["line_opacity"].compute_asynth_from_transp(), g["_add_fort19_asynth"](),
                g["_compute_hydrogen_line_opacity"](),
                g["solve_lte_spectrum"](), rad.solve_josh_flux(), g["bfudge"].compute_bfudge_and_slinec()]

    op.run_synthesis = run_synthesis
    yield op, calls
    for name in [n for n in sys.modules if n == PKG or n.startswith(PKG + ".")]:
        del sys.modules[name]


def test_install_reaches_every_engine_call_site(ref_layout):
    from kurucz_py_joss_submission_amd import integration

    op, calls = ref_layout
    impl = types.SimpleNamespace()
    for name in {b[2] for b in integration.BINDINGS}:
        setattr(impl, name, (lambda n: (lambda *a, **k: calls.append(("gpu", n)) or n))(name))
    assert op.run_synthesis() == ["continuum", "transp", "asynth", "fort19", "hydrogen", "lte_spectrum", "josh",
                                  "bfudge"]
    assert all(who == "ref" for who, _ in calls)
    calls.clear()
    saved = integration.install(PKG, impl)
    assert len(saved) == len(integration.BINDINGS)
    op.run_synthesis()
    assert calls == [("gpu", "build_depth_continuum"), ("gpu", "compute_transp"),
                     ("gpu", "compute_asynth_from_transp"), ("gpu", "add_fort19_asynth"),
                     ("gpu", "compute_hydrogen_line_opacity"),
                     ("gpu", "solve_lte_spectrum"),
                     ("gpu", "solve_josh_flux"), ("gpu", "compute_bfudge_and_slinec")]
    calls.clear()
    integration.uninstall(saved)
    op.run_synthesis()
    assert all(who == "ref" for who, _ in calls) and len(calls) == 8


def test_default_impl_is_the_drop_in_module(ref_layout):
    from kurucz_py_joss_submission_amd import integration, physics

    saved = integration.install(PKG)
    try:
        lo = sys.modules[f"{PKG}.physics.line_opacity"]
        assert lo.compute_transp is physics.compute_transp
        assert sys.modules[f"{PKG}.engine.opacity"].solve_lte_spectrum is physics.solve_lte_spectrum
    finally:
        integration.uninstall(saved)


def test_install_from_env_needs_the_switch(ref_layout, monkeypatch):
    from kurucz_py_joss_submission_amd import integration

    monkeypatch.delenv("SYNTHE_BACKEND", raising=False)
    assert integration.install_from_env(PKG) is None
    monkeypatch.setenv("SYNTHE_BACKEND", "mi355x")
    saved = integration.install_from_env(PKG)
    assert saved and len(saved) == len(integration.BINDINGS)
    integration.uninstall(saved)
