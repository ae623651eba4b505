"""Thread-safety contract (numbagg/test/test_multithreading.py:1-14): every
public function, called once serially and then 5 times from a 2-worker
ThreadPoolExecutor, returns the serial result each time.

Bit-exact functions, the rolling/EWMA/fill kernels and the plain reductions
(no cross-workgroup float merges in arbitrary order) must repeat their bits;
grouped float reductions merge per-workgroup float64 partials with atomics, so
their results may differ in the last place between runs: float32 within 1 ULP,
float64 within 1e-12 relative.
"""

from __future__ import annotations

from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ATOMIC_FLOAT = {"group_nansum", "group_nanmean", "group_nanvar", "group_nanstd", "group_nanprod",
                "group_nansum_of_squares"}


def _cases():
    import numbagg_amd as nb

    rng = np.random.default_rng(0)
    x = rng.standard_normal((3, 20_000))
    x[rng.random(x.shape) < 0.1] = np.nan
    y = rng.standard_normal((3, 20_000))
    lab = rng.integers(-1, 30, 20_000)
    cases = []
    xm = np.ascontiguousarray(x.T[:2000])  # (obs, vars) for the moving matrices
    for f in nb.MOVE_FUNCS:
        if f.__name__.endswith("matrix"):
            cases.append((f.__name__, lambda f=f: f(xm, window=50, min_count=5)))
            continue
        args = (x, y) if f.__name__ in ("move_cov", "move_corr") else (x,)
        cases.append((f.__name__, lambda f=f, args=args: f(*args, window=50, min_count=5)))
    for f in nb.MOVE_EXP_FUNCS:
        args = (x, y) if f.__name__ in ("move_exp_nancov", "move_exp_nancorr") else (x,)
        cases.append((f.__name__, lambda f=f, args=args: f(*args, alpha=0.1)))
    for f in nb.MOVE_EXP_MATRIX_FUNCS:
        cases.append((f.__name__, lambda f=f: f(xm, alpha=0.1)))
    for f in nb.MATRIX_FUNCS:
        cases.append((f.__name__, lambda f=f: f(x)))
    for f in nb.GROUPED_FUNCS:
        cases.append((f.__name__, lambda f=f: f(x, lab, num_labels=30, axis=-1)))
    for f in nb.OTHER_FUNCS:
        cases.append((f.__name__, lambda f=f: f(x, limit=7)))
    for f in nb.AGGREGATION_FUNCS:
        cases.append((f.__name__, lambda f=f: f(x, axis=-1)))
    cases.append(("factorize", lambda: nb.factorize(lab)[0]))
    return cases


def _same(name, a, b):
    a, b = np.asarray(a), np.asarray(b)
    if name in ATOMIC_FLOAT and a.dtype.kind == "f":
        # float64 sums merged in any order: within the float64 parity bar of
        # the grouped tests (1e-12 relative); float32: one rounding, 1 ULP
        assert np.array_equal(np.isnan(a), np.isnan(b)), name
        ok = ~np.isnan(a)
        if a.dtype == np.float32:
            bar = np.spacing(np.abs(a[ok]).astype(a.dtype))
        else:
            bar = 1e-12 * np.maximum(np.abs(a[ok]), 1e-300)
        assert np.all(np.abs(a[ok] - b[ok]) <= bar), name
    else:
        assert a.dtype == b.dtype and np.array_equal(a.view(np.uint8), b.view(np.uint8)), name


def test_every_function_from_two_threads(gpu_lib):
    for name, call in _cases():
        want = call()
        with ThreadPoolExecutor(max_workers=2) as ex:
            futs = [ex.submit(call) for _ in range(5)]
            got = [f.result() for f in futs]
        assert len(got) == 5
        for g in got:
            _same(name, g, want)


def test_interleaved_functions_from_four_threads(gpu_lib):
    """All functions at once from 4 threads: per-call workspaces and per-device
    caches stay separate."""
    cases = _cases()
    want = {name: call() for name, call in cases}
    with ThreadPoolExecutor(max_workers=4) as ex:
        futs = [(name, ex.submit(call)) for _ in range(2) for name, call in cases]
        for name, f in futs:
            _same(name, f.result(), want[name])
