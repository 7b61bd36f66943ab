"""C-ABI checks that need no GPU: the library loads, exports every function
include/inca_hip.h declares, and its host-side tables are right."""
import ctypes
import os
import re

import numpy as np
import pytest

from inca_parallel_amd import capi
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SIZES = [6, 14, 26, 38, 50, 74, 86, 110, 146, 170, 194, 230, 266, 302, 350, 434, 590, 770,
         974, 1202, 1454, 1730, 2030, 2354, 2702, 3074, 3470, 3890, 4334, 4802, 5294, 5810]
ORDERS = [3, 5, 7, 9, 11, 13, 15, 17, 19, 21, 23, 25, 27, 29, 31, 35, 41, 47, 53, 59, 65, 71,
          77, 83, 89, 95, 101, 107, 113, 119, 125, 131]


def declared_functions():
    txt = open(os.path.join(ROOT, "include", "inca_hip.h")).read()
    return sorted(set(re.findall(r"\b(inca_\w+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(capi.LIB_PATH)
    decl = declared_functions()
    assert len(decl) >= 20
    for name in decl:
        assert hasattr(lib, name), name
    assert sorted(capi.SYMBOLS) == decl


@pytest.mark.parametrize("n,order", list(zip(SIZES, ORDERS)))
def test_lebedev_rules(n, order):
    a = capi.lebedev(n)
    assert abs(a[:, 3].sum() - 1.0) < 1e-13
    assert np.abs(np.linalg.norm(a[:, :3], axis=1) - 1.0).max() < 1e-14
    # same point set as scipy's independent copy of the Lebedev-Laikov tables
    b = O.lebedev_unit(n)
    ka = np.lexsort(np.round(a, 12).T[::-1])
    kb = np.lexsort(np.round(b, 12).T[::-1])
    assert np.abs(a[ka] - b[kb]).max() < 1e-15
    # exact for monomials x^i y^j z^k with i+j+k <= order (sphere averages)
    rng = np.random.default_rng(n)
    for _ in range(6):
        i, j, k = (2 * int(v) for v in rng.integers(0, order // 6 + 1, size=3))
        if i + j + k > order:
            continue
        exact = _sphere_avg(i, j, k)
        got = np.sum(a[:, 3] * a[:, 0] ** i * a[:, 1] ** j * a[:, 2] ** k)
        assert abs(got - exact) < 1e-13


def _sphere_avg(i, j, k):
    """Average of x^i y^j z^k over the unit sphere (even exponents)."""
    from math import gamma
    return (gamma((i + 1) / 2) * gamma((j + 1) / 2) * gamma((k + 1) / 2) /
            gamma((i + j + k + 3) / 2)) / (2 * np.pi)


def test_unsupported_lebedev_size():
    with pytest.raises(capi.IncaError):
        capi.lebedev(100)


def test_gauleg():
    for n in (1, 3, 50, 200):
        t, w = capi.gauleg(n)
        t2, w2 = np.polynomial.legendre.leggauss(n)
        assert np.abs(t - t2).max() < 1e-14 and np.abs(w - w2).max() < 1e-13


def test_create_fails_cleanly_without_gpu():
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    with pytest.raises(capi.IncaError):
        capi.Engine(0)
