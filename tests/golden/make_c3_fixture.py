"""Write tests/golden/c3_oracle_sample.npz: the oracle's spectra for 16
instances of the batched configuration, sampled.

Configuration: 256 atmospheres (Teff ~ U(3500, 40000) K, log g ~ U(0, 5),
64-80 layers, [M/H] ~ U(-2, 0.5)), each with its own window (start ~
U(300, 1700) nm, width ~ U(5, 50) nm, R = 300000), all drawing on the shared
1,000,000-line list of the full-window workload (synthetic.batch_draws,
seed 3).  The 16 instances kept are the extremes of every draw (hottest,
coolest, lowest and highest log g, fewest and most layers, lowest and highest
[M/H], bluest and reddest window, narrowest and widest window) and, to fill
16, the coolest of the others (their spectra carry the most line opacity).
For each, the oracle (C restatement of the reference
hot path, reading the window's GFALL text with its own parser) runs the whole
spectrum; the fixture keeps ~200 evenly spaced wavelengths' flux and
continuum and the exact counts (line x depth x lambda evaluations, wing
pairs, JOSH iterations).

    python tests/golden/make_c3_fixture.py
"""

from __future__ import annotations

import os
import sys
import tempfile
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))

N_SAMPLES = 200


def chosen(draws) -> list:
    d = np.array(draws)
    width = d[:, 5] - d[:, 4]
    picks = [int(np.argmax(d[:, 0])), int(np.argmin(d[:, 0])), int(np.argmin(d[:, 1])), int(np.argmax(d[:, 1])),
             int(np.argmin(d[:, 2])), int(np.argmax(d[:, 2])), int(np.argmin(d[:, 3])), int(np.argmax(d[:, 3])),
             int(np.argmin(d[:, 4])), int(np.argmax(d[:, 4])), int(np.argmin(width)), int(np.argmax(width))]
    out = list(dict.fromkeys(picks))
    for i in np.argsort(d[:, 0]).tolist():  # then the coolest others: the most line opacity
        if len(out) == 16:
            break
        if i not in out:
            out.append(int(i))
    return out


def main() -> None:
    import catalog as C
    import oracle
    from kurucz_py_joss_submission_amd import synthetic as S

    oracle.build()
    oracle.set_threads(os.cpu_count() or 1)
    wk = S.make_workload("c2", 4000.0, 5.0, 1_000_000, 300.0, 1800.0, seed=1)
    draws = S.batch_draws()
    idx = chosen(draws)
    rec = {k: [] for k in ("instance", "n_wavelengths", "e_line", "wing_pairs", "josh_iters")}
    samples = []
    for i in idx:
        teff, logg, nl, mh, w0, w1 = draws[i]
        t0 = time.perf_counter()
        atm = S.build_bundle(S.make_atm(teff, logg, nl, mh))
        with tempfile.TemporaryDirectory() as td:
            gfall = S.write_workload_gfall(wk, Path(td) / "lines.gfall", window=(w0, w1))
            cat = C.filter_by_range(C.read_gfall(gfall), w0, w1)
        r = oracle.synthesize(atm, cat, w0, w1, 300000.0, keep_stages=True)
        W = r.wavelength.size
        k = np.unique(np.linspace(0, W - 1, N_SAMPLES).round().astype(np.int64))
        samples.append((k, r.flux[k], r.continuum[k], r.stages["iters"][k]))
        rec["instance"].append(i)
        rec["n_wavelengths"].append(W)
        rec["e_line"].append(r.stats["e_line"])
        rec["wing_pairs"].append(r.stats["wing_pairs"])
        rec["josh_iters"].append(r.stats["josh_iters"])
        print(f"instance {i}: Teff {teff:.0f} log g {logg:.2f} D {nl} [M/H] {mh:+.2f} {w0:.1f}-{w1:.1f} nm "
              f"W {W} lines {len(cat)} E_line {r.stats['e_line']:.4g} ({time.perf_counter() - t0:.1f} s)",
              flush=True)
    offs = np.cumsum([0] + [len(s[0]) for s in samples])
    np.savez_compressed(
        Path(__file__).with_name("c3_oracle_sample.npz"),
        **{k: np.array(v) for k, v in rec.items()},
        offsets=offs, index=np.concatenate([s[0] for s in samples]),
        flux=np.concatenate([s[1] for s in samples]), continuum=np.concatenate([s[2] for s in samples]),
        iters=np.concatenate([s[3] for s in samples]))


if __name__ == "__main__":
    main()
