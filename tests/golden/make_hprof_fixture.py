"""Write tests/golden/hprof_reference.npz: values of the reference's HPROF4
hydrogen profile (synthe_py/physics/profiles/hydrogen.py:1541
hydrogen_line_profile), its Stark function SOFBETA (:1451) and VCSE1F (:1408)
on a seeded spread of inputs.

The profile's composite logic (width selection, the Y1 weights, the Stark
gamma, the Lyman-alpha cutoffs) has no closed form to pin the oracle's
restatement against, so these vectors come from the reference itself, read
in this container from /root/reference: the three modules the function needs
(physics/tables.py, physics/populations.py, physics/profiles/hydrogen.py)
are loaded from their files, without the package __init__ files, which pull
in modules that need numba.  Only inputs and outputs are committed.

    python tests/golden/make_hprof_fixture.py [REFERENCE_ROOT]
"""

from __future__ import annotations

import importlib.util
import sys
import types
from pathlib import Path

import numpy as np

OUT = Path(__file__).with_name("hprof_reference.npz")


def _load(root: Path):
    pkg = root / "synthe_py"
    for name, path in (("synthe_py", pkg), ("synthe_py.physics", pkg / "physics"),
                       ("synthe_py.physics.profiles", pkg / "physics" / "profiles")):
        mod = types.ModuleType(name)
        mod.__path__ = [str(path)]
        sys.modules[name] = mod
    mods = {}
    for name, rel in (("synthe_py.physics.tables", "physics/tables.py"),
                      ("synthe_py.physics.populations", "physics/populations.py"),
                      ("synthe_py.physics.profiles.hydrogen", "physics/profiles/hydrogen.py")):
        spec = importlib.util.spec_from_file_location(name, pkg / rel)
        m = importlib.util.module_from_spec(spec)
        sys.modules[name] = m
        spec.loader.exec_module(m)
        mods[name.rsplit(".", 1)[-1]] = m
    return mods


def main(root: str = "/root/reference") -> None:
    mods = _load(Path(root))
    H, P = mods["hydrogen"], mods["populations"]
    rng = np.random.default_rng(11)
    # depth states over a cool dwarf to a hot giant
    states, sfields = [], []
    for _ in range(12):
        T = float(rng.uniform(3500.0, 30000.0))
        xne = float(10 ** rng.uniform(9.0, 16.0))
        he1 = float(10 ** rng.uniform(12.0, 16.0))
        h2 = float(10 ** rng.uniform(0.0, 15.0)) if rng.uniform() < 0.7 else 0.0
        xnfph = np.array([10 ** rng.uniform(10.0, 17.0), 10 ** rng.uniform(8.0, 15.0)])
        dopph = float(rng.uniform(5e-6, 5e-5))
        hs = P._hydrogen_state(T, xne, he1, h2, xnfph, dopph)
        ds = P.DepthState(boltzmann_factor=np.zeros(1), doppler_width=np.zeros(1), turbulence_width=0.0,
                          electron_density=xne, temperature=T, continuum_opacity=np.zeros(1), hckt=1.4388 / T,
                          txnxn=0.0, hydrogen=hs)
        states.append(ds)
        sfields.append([T, xne, he1, h2, xnfph[0], xnfph[1], dopph])
    # transitions: Lyman, Balmer, Paschen, Brackett, Pfund and high members
    pairs = [(1, 2), (1, 3), (1, 4), (1, 5), (1, 8), (2, 3), (2, 4), (2, 5), (2, 6), (2, 12), (2, 30),
             (3, 4), (3, 5), (3, 9), (3, 20), (4, 5), (4, 7), (4, 12), (5, 6), (5, 9), (6, 8), (7, 11)]
    rows = []
    for n, m in pairs:
        for si, ds in enumerate(states):
            # offsets from the line centre: the core, the near wing, the far wing
            for dl in np.concatenate([[0.0], rng.uniform(-0.01, 0.01, 3), rng.uniform(-1.0, 1.0, 3),
                                      rng.uniform(-40.0, 40.0, 3)]):
                rows.append((n, m, si, float(dl), H.hydrogen_line_profile(n, m, ds, float(dl))))
    sof = []
    for _ in range(400):
        beta = float(10 ** rng.uniform(-2.0, 3.5))
        pp = float(rng.uniform(0.0, 1.2))
        n = int(rng.integers(1, 6))
        m = n + int(rng.integers(1, 6))
        sof.append((beta, pp, n, m, H.sofbeta(beta, pp, n, m)))
    for beta in (1.0, 8.0, 10.0, 25.12, 25.13, 500.0, 500.5):
        sof.append((beta, 0.3, 2, 3, H.sofbeta(beta, 0.3, 2, 3)))
    xs = np.concatenate([10 ** rng.uniform(-4.0, 1.6, 300), [0.005, 0.01, 0.0100001, 1.0, 1.0000001, 30.0, 31.0]])
    vc = [(float(x), H.vcse1f(float(x))) for x in xs]
    np.savez_compressed(OUT, state_inputs=np.array(sfields), profile=np.array(rows), sofbeta=np.array(sof),
                        vcse1f=np.array(vc))
    print(f"wrote {OUT}: {len(rows)} profile values, {len(sof)} SOFBETA, {len(vc)} VCSE1F")


if __name__ == "__main__":
    main(*sys.argv[1:])
