"""Command line mirroring synthe_py/cli.py:15-150 (same positional arguments and flags).

    python -m kurucz_py_joss_submission_amd.cli MODEL.npz LINES.gfall --wl-start 500 --wl-end 510

Flags the reference parses but never uses downstream (--linout, --nlte,
--scat-iterations, --scat-tol) are accepted the same way.
"""

from __future__ import annotations

import argparse
import logging
from pathlib import Path
from typing import Optional

from .engine import SynthesisConfig, run_synthesis


def build_parser() -> argparse.ArgumentParser:
    p = argparse.ArgumentParser(description="synthe hot path on MI355X")
    p.add_argument("model", type=Path, help="Atmosphere bundle (.npz), or .atm with an .npz beside it")
    p.add_argument("atomic", type=Path, help="Atomic line catalog (GFALL fixed-width, or parsed .npz)")
    p.add_argument("--npz", type=Path, default=None)
    p.add_argument("--spec", type=Path, default=Path("spectrum.spec"))
    p.add_argument("--wl-start", type=float, default=300.0)
    p.add_argument("--wl-end", type=float, default=1800.0)
    p.add_argument("--resolution", type=float, default=300_000.0)
    p.add_argument("--no-vacuum", action="store_true")
    p.add_argument("--cutoff", type=float, default=1e-3)
    p.add_argument("--linout", type=int, default=30)
    p.add_argument("--cache", type=Path, default=None)
    p.add_argument("--allow-tfort-runtime", action="store_true")
    p.add_argument("--n-workers", type=int, default=None)
    p.add_argument("--nlte", action="store_true")
    p.add_argument("--scat-iterations", type=int, default=8)
    p.add_argument("--scat-tol", type=float, default=1e-3)
    p.add_argument("--rhoxj", type=float, default=0.0)
    p.add_argument("--log-level", type=str, default="INFO", choices=["DEBUG", "INFO", "WARNING", "ERROR"])
    p.add_argument("--diagnostics", type=Path, default=None)
    p.add_argument("--microturb", type=float, default=0.0)
    return p


def main(argv: Optional[list] = None) -> int:
    args = build_parser().parse_args(argv)
    cfg = SynthesisConfig.from_cli(
        spec_path=args.spec, diagnostics_path=args.diagnostics, atmosphere_path=args.model,
        atomic_catalog=args.atomic, wl_start=args.wl_start, wl_end=args.wl_end,
        resolution=args.resolution, velocity_microturb=args.microturb, vacuum=not args.no_vacuum,
        cutoff=args.cutoff, linout=args.linout, nlte=args.nlte,
        scattering_iterations=args.scat_iterations, scattering_tolerance=args.scat_tol,
        rhoxj_scale=args.rhoxj, npz_path=args.npz, n_workers=args.n_workers,
        allow_tfort_runtime=args.allow_tfort_runtime)
    if args.cache:
        cfg.line_data.cache_directory = args.cache
    cfg.log_level = args.log_level
    logging.basicConfig(level=getattr(logging, cfg.log_level), format="%(levelname)s %(name)s: %(message)s")
    run_synthesis(cfg)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
