"""CPU-side checks of the product library: it loads, exports every symbol
include/artea_hip.h declares, and its host-only parts behave (workload
generator determinism).  No GPU calls."""
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "artea_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(artea_[a-z_0-9]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    from artea_amd import _lib

    l = _lib.lib()
    missing = [s for s in declared_symbols() if not hasattr(l, s)]
    assert not missing, missing
    assert set(declared_symbols()) <= set(_lib.SIGNATURES), "ctypes signatures missing for some symbols"
    assert b"gfx950" in l.artea_version()


def test_generator_is_deterministic_and_windowed():
    import artea_amd as aa

    a = aa.generate(0, 20261204, 0, 500, 128, nthreads=1)
    b = aa.generate(0, 20261204, 0, 500, 128, nthreads=4)
    assert np.array_equal(a, b)
    c = aa.generate(0, 20261204, 0, 200, 128, first=300)
    assert np.array_equal(a[300:], c)
    assert a.min() >= 0 and a.max() <= 255 and np.all(a == np.rint(a))
    q = aa.generate(0, 20261204, 1, 10, 128)
    assert not np.array_equal(q, a[:10])
    x = aa.generate(1, 5, 0, 100, 100)
    assert x.shape == (100, 112) and np.all(x[:, 100:] == 0)
    assert np.allclose(np.linalg.norm(x, axis=1), 1, atol=1e-5)
    y = aa.generate(2, 5, 0, 100, 96)
    assert np.allclose(np.linalg.norm(y, axis=1), 1, atol=1e-5)


def test_generator_mixture_statistics():
    import artea_amd as aa

    a = aa.generate(0, 7, 0, 20000, 128).astype(np.float64)
    # centres U[0,128] and sigma 12: mean 64, per-coordinate variance 128^2/12 + 144
    assert abs(a.mean() - 64) < 2
    assert abs(a.var() - (128 ** 2 / 12 + 144)) < 60


def test_error_path_without_gpu_is_reported():
    import artea_amd as aa

    try:
        aa.generate(5, 0, 0, 1, 16)
    except aa.ArteaError as e:
        assert "kind" in str(e)
    else:
        raise AssertionError("expected an error")
