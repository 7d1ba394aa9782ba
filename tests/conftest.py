"""Shared pytest configuration.

Tests marked ``gpu`` run the HIP library on an MI355X and compare it with the
CPU oracle (``oracle/``); everything else runs on the CPU.
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def rel_l2(got_re, got_im, ref_re, ref_im) -> float:
    """Normwise relative error of (re, im) vectors (benches/support/mod.rs:114-123)."""
    num = np.sum((np.asarray(got_re, np.float64) - ref_re) ** 2
                 + (np.asarray(got_im, np.float64) - ref_im) ** 2)
    den = np.sum(np.asarray(ref_re) ** 2 + np.asarray(ref_im) ** 2)
    return float(np.sqrt(num / den))


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def shk():
    import shrinkers_amd
    from shrinkers_amd import _lib
    _lib.load()
    return shrinkers_amd
