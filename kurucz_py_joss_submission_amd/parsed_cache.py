"""Parsed GFALL cache: the parsed line catalog kept as arrays next to the
GFALL file, so later runs skip the text parse (synthe_py/io/lines/parsed_cache.py).

Same contract as the reference's module:

* the cache is window-independent: one file per catalog, reused by every
  wavelength window (parsed_cache.py:1-6);
* it is keyed and validated by the source's fingerprint -- resolved path,
  size, mtime_ns -- plus a schema and logic version; the key is the first 24
  hex digits of the SHA-256 of that payload (parsed_cache.py:37-60); a JSON
  manifest beside the arrays must equal the payload, else the cache is stale
  and rebuilt (parsed_cache.py:267-284);
* default directory ``<catalog dir>/.py_line_cache`` (parsed_cache.py:33-34),
  files ``parsed_gfall_<key>.npz`` / ``.json`` (parsed_cache.py:63-75);
* an unreadable cache is rebuilt, not an error (parsed_cache.py:287-303);
* the window form returns the records inside the DELLIM margins of
  [wlbeg, wlend] (parsed_cache.py:198-225, 306-323), the same selection as
  ``lines.filter_by_range``.

The arrays are this package's LineCatalog fields (a structure of arrays), so
saving and loading are plain array I/O, with no per-record objects; the
window form reads only the columns it needs to select before gathering the
rest.  The schema string differs from the reference's, so the two caches never
take each other's files even in one directory.
"""

from __future__ import annotations

import hashlib
import json
import zipfile
from dataclasses import dataclass
from pathlib import Path
from typing import Any, Dict, Optional

import numpy as np

from . import lines as L

PARSED_CACHE_SCHEMA = "synthe-hip-soa-1"
PARSED_CACHE_LOGIC_VERSION = 2
# what reading a damaged or foreign cache file can raise
_UNREADABLE = (OSError, ValueError, KeyError, TypeError, EOFError, zipfile.BadZipFile)


@dataclass(frozen=True)
class ParsedCachePaths:
    """Paths of one catalog's cache files."""

    npz: Path
    manifest: Path


def _default_cache_dir(catalog_path: Path) -> Path:
    return Path(catalog_path).parent / ".py_line_cache"


def parsed_cache_key_payload(catalog_path: Path) -> Dict[str, Any]:
    """The fingerprint the cache is keyed and validated by."""
    p = Path(catalog_path)
    st = p.stat()
    return {"schema": PARSED_CACHE_SCHEMA, "logic_version": PARSED_CACHE_LOGIC_VERSION,
            "source": str(p.resolve()), "size": int(st.st_size),
            "mtime_ns": int(getattr(st, "st_mtime_ns", int(st.st_mtime * 1e9)))}


def _parsed_key(catalog_path: Path) -> str:
    payload = parsed_cache_key_payload(catalog_path)
    return hashlib.sha256(json.dumps(payload, sort_keys=True).encode("utf-8")).hexdigest()[:24]


def cache_paths(catalog_path: Path, cache_directory: Optional[Path] = None) -> ParsedCachePaths:
    d = Path(cache_directory) if cache_directory is not None else _default_cache_dir(catalog_path)
    key = _parsed_key(catalog_path)
    return ParsedCachePaths(npz=d / f"parsed_gfall_{key}.npz", manifest=d / f"parsed_gfall_{key}.json")


def _manifest_valid(paths: ParsedCachePaths, catalog_path: Path) -> bool:
    if not (paths.npz.exists() and paths.manifest.exists()):
        return False
    try:
        return json.loads(paths.manifest.read_text(encoding="utf-8")) == parsed_cache_key_payload(catalog_path)
    except (OSError, ValueError):
        return False


def _write(paths: ParsedCachePaths, catalog_path: Path, catalog: L.LineCatalog) -> None:
    """Store the cache; a directory that cannot be written leaves the run
    uncached rather than failing it."""
    try:
        paths.npz.parent.mkdir(parents=True, exist_ok=True)
        # write then rename, so a reader never sees half a file
        tmp = paths.npz.with_name(paths.npz.stem + ".partial.npz")
        catalog.save_npz(tmp)
        tmp.replace(paths.npz)
        paths.manifest.write_text(json.dumps(parsed_cache_key_payload(catalog_path), indent=2) + "\n",
                                  encoding="utf-8")
    except OSError:
        pass


def load_or_build_parsed_catalog(catalog_path: Path, cache_directory: Optional[Path] = None) -> L.LineCatalog:
    """The parsed catalog of a GFALL file, from the cache when it is current,
    else parsed (lines.load_catalog) and cached."""
    catalog_path = Path(catalog_path)
    paths = cache_paths(catalog_path, cache_directory)
    if _manifest_valid(paths, catalog_path):
        try:
            return L.LineCatalog.load_npz(paths.npz)
        except _UNREADABLE:
            pass
    catalog = L.load_catalog(catalog_path)
    _write(paths, catalog_path, catalog)
    return catalog


def load_or_build_parsed_catalog_window(catalog_path: Path, wlbeg: float, wlend: float,
                                        cache_directory: Optional[Path] = None) -> L.LineCatalog:
    """The records of [wlbeg, wlend] plus their DELLIM margins, read from the
    cache (building it first when it is missing or stale)."""
    catalog_path = Path(catalog_path)
    paths = cache_paths(catalog_path, cache_directory)
    if _manifest_valid(paths, catalog_path):
        try:
            with np.load(paths.npz, allow_pickle=False) as z:
                wl = z["wavelength"]
                if wl.size == 0:
                    return L.LineCatalog.load_npz(paths.npz)
                margin = L.line_margin(z["line_types"], z["atomic_number"], wlbeg,
                                       z["line_size"] if "line_size" in z.files else None,
                                       z["code"] if "code" in z.files else None)
                keep = np.nonzero((wl >= wlbeg - margin) & (wl <= wlend + margin))[0]
                return L.LineCatalog(**{k: z[k][keep] for k in z.files})
        except _UNREADABLE:
            pass
    return L.filter_by_range(load_or_build_parsed_catalog(catalog_path, cache_directory), wlbeg, wlend)


def load_manifest(catalog_path: Path, cache_directory: Optional[Path] = None) -> Optional[Dict[str, Any]]:
    """The cache manifest of a catalog, if one is present."""
    paths = cache_paths(catalog_path, cache_directory)
    if not paths.manifest.exists():
        return None
    try:
        return json.loads(paths.manifest.read_text(encoding="utf-8"))
    except (OSError, ValueError):
        return None
