"""Shared test configuration.

Markers: ``gpu`` (needs a ROCm GPU and the built HIP library; parity tests call
through the C ABI) and ``slow``. CPU-only runs use ``-m "not gpu"``.
"""

from __future__ import annotations

import pathlib
import sys

import numpy as np
import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

GOLDEN = ROOT / "tests" / "golden" / "pandas_fixtures.npz"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU and the built HIP library")
    config.addinivalue_line("markers", "slow: long-running test")


def pytest_addoption(parser):
    parser.addoption("--skip-slow", action="store_true", default=False, help="skip slow tests")


def pytest_collection_modifyitems(config, items):
    if config.getoption("--skip-slow"):
        skip = pytest.mark.skip(reason="--skip-slow")
        for item in items:
            if "slow" in item.keywords:
                item.add_marker(skip)


@pytest.fixture(scope="session")
def golden():
    with np.load(GOLDEN, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="session")
def gpu_lib():
    """Build (if stale) and load the HIP library; skip if no GPU is visible."""
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no ROCm GPU visible")
    from numbagg_amd import _build, _lib

    if _build.library_is_stale():
        _build.build_library()
    lib = _lib.load()
    _lib.call("nbg_reset_spin_timeouts")
    yield lib
    assert _lib.spin_timeouts() == 0, "a bounded device-side wait expired"


def fixture_array(shape, seed=0):
    """The reference's default test array (numbagg/test/conftest.py:589-592)."""
    a = np.random.RandomState(seed).rand(*shape)
    return np.where(a > 0.1, a, np.nan)
