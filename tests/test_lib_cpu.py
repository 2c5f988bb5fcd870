"""CPU checks of the C-ABI library: it loads, exports every declared entry
point, and its host-only functions agree with the oracle.  No kernel runs."""
import os
import re

import numpy as np
import pytest

from critic2_amd import _lib
from critic2_amd.synth import lattice_from_params
from oracle import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    txt = open(os.path.join(ROOT, "include", "c2gpu.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(c2gpu_\w+)\s*\(", txt, re.M)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = declared()
    assert len(names) >= 13
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(_lib.EXPORTS)


@pytest.mark.parametrize("x2c", [np.diag([10.0, 12.0, 13.0]), lattice_from_params(20, 22, 25, 75, 100, 110),
                                 lattice_from_params(40, 40, 40, 71.5, 108.2, 93.3)])
def test_metrics_bitwise_equal_to_oracle(x2c):
    n = (48, 64, 80)
    a = _lib.bader_metrics(x2c, n)
    b = oracle.metrics(x2c, n)
    for u, v in zip(a, b):
        assert np.array_equal(u, v)


def test_context_without_gpu_fails_loudly():
    """On a host without a HIP device the library refuses to run (status
    C2GPU_ERR_HIP), it never computes on the CPU instead."""
    try:
        c = _lib.Context(0)
    except _lib.C2GpuError as e:
        assert e.code == 2
        return
    c.close()
    pytest.skip("a GPU is present")
