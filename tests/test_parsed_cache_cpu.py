"""Parsed GFALL cache (parsed_cache.py, after synthe_py/io/lines/parsed_cache.py):
the cached catalog equals a fresh parse field by field, the window form equals
filter_by_range of the full parse, a changed or corrupt source/cache is
rebuilt, and a current cache is read without parsing."""

from __future__ import annotations

import os

import numpy as np
import pytest


def _same(a, b):
    for k in a.__dataclass_fields__:
        x, y = getattr(a, k), getattr(b, k)
        if x is None or y is None:
            assert x is None and y is None, k
            continue
        assert x.dtype == y.dtype, k
        assert np.array_equal(x, y), k


@pytest.fixture()
def gfall(tmp_path):
    from kurucz_py_joss_submission_amd import synthetic as S

    path = tmp_path / "gfall.dat"
    S.write_line_list(path, S.synthetic_lines(20_000, 500.0, 530.0, seed=3))
    return path


def test_cache_round_trip_and_reuse(gfall, monkeypatch):
    from kurucz_py_joss_submission_amd import lines as L
    from kurucz_py_joss_submission_amd import parsed_cache as PC

    ref = L.load_catalog(gfall)
    c1 = PC.load_or_build_parsed_catalog(gfall)
    paths = PC.cache_paths(gfall)
    assert paths.npz.parent == gfall.parent / ".py_line_cache"
    assert paths.npz.exists() and PC.load_manifest(gfall) == PC.parsed_cache_key_payload(gfall)
    _same(c1, ref)

    def no_parse(*a, **k):
        raise AssertionError("parsed although the cache is current")

    monkeypatch.setattr(L, "load_catalog", no_parse)
    _same(PC.load_or_build_parsed_catalog(gfall), ref)
    for a, b in ((500.0, 505.0), (512.3, 512.4), (529.0, 540.0), (100.0, 200.0)):
        _same(PC.load_or_build_parsed_catalog_window(gfall, a, b), L.filter_by_range(ref, a, b))


def test_window_without_cache_and_custom_dir(gfall, tmp_path):
    from kurucz_py_joss_submission_amd import lines as L
    from kurucz_py_joss_submission_amd import parsed_cache as PC

    d = tmp_path / "elsewhere"
    w = PC.load_or_build_parsed_catalog_window(gfall, 505.0, 507.5, cache_directory=d)
    _same(w, L.filter_by_range(L.load_catalog(gfall), 505.0, 507.5))
    assert PC.cache_paths(gfall, d).npz.exists()
    assert not PC.cache_paths(gfall).npz.exists()


def test_stale_and_corrupt_caches_are_rebuilt(gfall):
    from kurucz_py_joss_submission_amd import lines as L
    from kurucz_py_joss_submission_amd import parsed_cache as PC
    from kurucz_py_joss_submission_amd import synthetic as S

    PC.load_or_build_parsed_catalog(gfall)
    old = PC.cache_paths(gfall)
    # a new source: new fingerprint, new key, new contents
    S.write_line_list(gfall, S.synthetic_lines(5_000, 600.0, 610.0, seed=4))
    st = os.stat(gfall)
    os.utime(gfall, ns=(st.st_atime_ns, st.st_mtime_ns + 1_000_000_000))
    new = PC.cache_paths(gfall)
    assert new.npz != old.npz
    _same(PC.load_or_build_parsed_catalog(gfall), L.load_catalog(gfall))
    # a corrupt cache file with a valid manifest: not a zip, then a truncated one
    new.npz.write_bytes(b"not an npz")
    _same(PC.load_or_build_parsed_catalog(gfall), L.load_catalog(gfall))
    data = new.npz.read_bytes()
    new.npz.write_bytes(data[: len(data) // 2])
    _same(PC.load_or_build_parsed_catalog_window(gfall, 601.0, 602.0),
          L.filter_by_range(L.load_catalog(gfall), 601.0, 602.0))
    new.npz.write_bytes(data[: len(data) // 2])
    _same(PC.load_or_build_parsed_catalog(gfall), L.load_catalog(gfall))
    # a manifest that does not match
    new.manifest.write_text("{}")
    _same(PC.load_or_build_parsed_catalog_window(gfall, 601.0, 602.0),
          L.filter_by_range(L.load_catalog(gfall), 601.0, 602.0))
    assert PC.load_manifest(gfall) == PC.parsed_cache_key_payload(gfall)


def test_empty_catalog(tmp_path):
    from kurucz_py_joss_submission_amd import lines as L
    from kurucz_py_joss_submission_amd import parsed_cache as PC

    p = tmp_path / "empty.dat"
    p.write_text("")
    c = PC.load_or_build_parsed_catalog(p)
    assert len(c) == 0
    assert len(PC.load_or_build_parsed_catalog_window(p, 500.0, 510.0)) == 0
    _same(PC.load_or_build_parsed_catalog(p), L.load_catalog(p))
