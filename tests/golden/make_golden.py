"""Generate the golden fixtures in tests/golden/ (run once; outputs are committed).

The reference numbagg checks its hot path against pandas on fixed inputs
(numbagg/test/conftest.py:230-550, test_moving.py:29-53, test_moving_exp.py:19-31,
test_funcs.py:38-53, test_grouped.py:137-166). This script evaluates the same
pandas expressions on the same deterministic inputs and stores inputs and
outputs as plain arrays (no pickles):

* array: np.random.RandomState(0).rand(3, 500) with values <= 0.1 set to NaN
  (conftest.py:581-592), and (1, 500) for the grouped comparisons;
* two-array functions use (a, a**2 + 1) (conftest.py:73-77);
* grouped labels: np.random.seed(0); randint(0, 12) (conftest.py:164-172).

Usage: python tests/golden/make_golden.py
"""

from __future__ import annotations

import pathlib
import warnings

import numpy as np
import pandas as pd

OUT = pathlib.Path(__file__).resolve().parent / "pandas_fixtures.npz"


def fixture_array(shape):
    a = np.random.RandomState(0).rand(*shape)
    return np.where(a > 0.1, a, np.nan)


def df_of(a):
    return pd.DataFrame(a).T if a.ndim == 2 else pd.DataFrame(a)


def main() -> None:
    res: dict[str, np.ndarray] = {}
    a = fixture_array((3, 500))
    res["array_3x500"] = a
    with warnings.catch_warnings():
        warnings.simplefilter("ignore", RuntimeWarning)
        a2 = a**2 + 1
    res["array2_3x500"] = a2

    # rolling: windows {10, 50} x min_count {None, 0, 1, 3, window}
    for window in (10, 50):
        for mc in (None, 0, 1, 3, "window"):
            m = window if mc == "window" else mc
            tag = f"w{window}_mc{'None' if mc is None else m}"
            r = df_of(a).rolling(window=window, min_periods=m)
            res[f"move_mean_{tag}"] = r.mean().T.values
            res[f"move_sum_{tag}"] = r.sum().T.values
            res[f"move_std_{tag}"] = r.std().T.values
            res[f"move_var_{tag}"] = r.var().T.values
            res[f"move_cov_{tag}"] = r.cov(df_of(a2)).T.values
            res[f"move_corr_{tag}"] = r.corr(df_of(a2)).T.values

    # exponentially weighted: alpha {0.5, 0.1}
    for alpha in (0.5, 0.1):
        tag = f"a{alpha}"
        e = df_of(a).ewm(alpha=alpha)
        res[f"move_exp_nanmean_{tag}"] = e.mean().T.values
        res[f"move_exp_nanvar_{tag}"] = e.var().T.values
        res[f"move_exp_nanstd_{tag}"] = e.std().T.values
        res[f"move_exp_nansum_{tag}"] = e.sum().T.values
        res[f"move_exp_nancount_{tag}"] = df_of(a).notnull().ewm(alpha=alpha).sum().T.values
        res[f"move_exp_nancorr_{tag}"] = e.corr(df_of(a2)).T.values
        res[f"move_exp_nancov_{tag}"] = e.cov(df_of(a2)).T.values

    # fill: limit {1, 3, None}
    for limit in (1, 3, None):
        res[f"ffill_l{limit}"] = df_of(a).ffill(limit=limit).T.values
        res[f"bfill_l{limit}"] = df_of(a).bfill(limit=limit).T.values

    # grouped: (1, 500) array, labels randint(0, 12) seeded with 0
    g = fixture_array((1, 500))
    res["group_array_1x500"] = g
    np.random.seed(0)
    labels = np.random.randint(0, 12, size=500)
    res["group_labels"] = labels
    codes, uniques = pd.factorize(labels, sort=True)
    res["group_codes"] = codes
    gb = df_of(g).groupby(labels)
    res["group_nansum"] = gb.sum().T.values
    res["group_nanmean"] = gb.mean().T.values
    res["group_nancount"] = gb.count().T.values
    res["group_nanfirst"] = gb.first().T.values
    res["group_nanlast"] = gb.last().T.values
    res["group_nanargmax"] = gb.idxmax().T.values.astype(np.float64)
    res["group_nanargmin"] = gb.idxmin().T.values.astype(np.float64)
    for ddof in (0, 1, 5):
        res[f"group_nanvar_ddof{ddof}"] = gb.var(ddof=ddof).T.values
        res[f"group_nanstd_ddof{ddof}"] = gb.std(ddof=ddof).T.values
    # the rest of the grouped family (conftest.py:444-487)
    res["group_nanmin"] = gb.min().T.values
    res["group_nanmax"] = gb.max().T.values
    res["group_nanprod"] = gb.prod().T.values
    res["group_nanany"] = gb.any().T.values.astype(np.float64)
    res["group_nanall"] = gb.all().T.values.astype(np.float64)
    res["group_nansum_of_squares"] = df_of(g).pipe(lambda x: x**2).groupby(labels).sum().T.values

    # test_grouped.py:67-135: labels rs.choice([-1, 0, .., 4], 200) and float64 /
    # int32 / bool values drawn from a RandomState(0) stream, compared with
    # pandas Series.groupby (NaN labels for -1)
    def series_funcs():
        return {
            "group_nanall": lambda x: x.all(),
            "group_nanany": lambda x: x.any(),
            "group_nancount": lambda x: x.count(),
            "group_nanfirst": lambda x: x.first(),
            "group_nanlast": lambda x: x.last(),
            "group_nanmax": lambda x: x.max(),
            "group_nanmean": lambda x: x.mean(),
            "group_nanmin": lambda x: x.min(),
            "group_nanprod": lambda x: x.prod(),
            "group_nanstd": lambda x: x.std(),
            "group_nansum": lambda x: x.sum(),
            "group_nanvar": lambda x: x.var(),
            "group_nansum_of_squares": lambda x: x.agg(lambda y: (y**2).sum()),
        }

    for dt in ("float64", "int32", "bool"):
        rs = np.random.RandomState(0)
        labels200 = rs.choice([-1, 0, 1, 2, 3, 4], size=200)
        if dt == "int32":
            vals = rs.randint(-100, 100, size=200)
        elif dt == "float64":
            vals = rs.rand(200)
            vals = np.where(vals > 0.1, vals, np.nan)
            vals = np.where(labels200 != 2, vals, np.nan)
        else:
            vals = rs.choice([True, False], size=200)
        res[f"series_{dt}_labels"] = labels200
        res[f"series_{dt}_values"] = vals
        plabels = np.where(labels200 >= 0, labels200, np.nan)
        for name, fn in series_funcs().items():
            with warnings.catch_warnings():
                warnings.simplefilter("ignore")
                res[f"series_{dt}_{name}"] = np.asarray(fn(pd.Series(vals).groupby(plabels)).values, dtype=np.float64)

    # matrix functions (conftest.py:108-156, :204-211): static on (vars, obs)
    # arrays, moving / exponential on the (obs, vars) transpose; pandas' pairwise
    # DataFrame statistics use pairwise-complete observations.
    for shape in ((3, 500), (20, 1000)):
        m = a if shape == (3, 500) else fixture_array(shape)
        tag = f"{shape[0]}x{shape[1]}"
        res[f"matrix_array_{tag}"] = m
        res[f"nancorrmatrix_{tag}"] = pd.DataFrame(m.T).corr().values
        res[f"nancovmatrix_{tag}"] = pd.DataFrame(m.T).cov().values
    mt = np.ascontiguousarray(a.T)  # (500, 3)
    for window, mc in ((20, None), (20, 3), (50, 1)):
        tag = f"w{window}_mc{mc}"
        r = pd.DataFrame(mt).rolling(window, min_periods=mc)
        res[f"move_corrmatrix_{tag}"] = r.corr().values.reshape(500, 3, 3)
        res[f"move_covmatrix_{tag}"] = r.cov().values.reshape(500, 3, 3)
    for alpha in (0.5, 0.1):
        e = pd.DataFrame(mt).ewm(alpha=alpha)
        res[f"move_exp_nancorrmatrix_a{alpha}"] = e.corr().values.reshape(500, 3, 3)
        res[f"move_exp_nancovmatrix_a{alpha}"] = e.cov().values.reshape(500, 3, 3)

    np.savez_compressed(OUT, **res)
    print(f"wrote {OUT} ({len(res)} arrays, pandas {pd.__version__})")


if __name__ == "__main__":
    main()
