"""CPU oracle for numbagg's hot path — TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
this package, and only as the checker (or the timed CPU baseline). The
numbagg_amd product never imports it.

Two layers:

* ``restated``: ctypes bindings of ``nbg_oracle.c``, a plain-C restatement of
  the reference kernels with numba's accumulator types and operation order
  (each C function cites the numbagg file:line it follows).
* ``exact``: extended-precision / exact oracles used to state float error
  bounds (long-double window and recurrence evaluation in C; NumPy closed forms
  for fill and grouped reductions).

All functions take (rows, n) NumPy arrays and return NumPy arrays.
"""

from __future__ import annotations

import ctypes
import math
import pathlib
import subprocess

import numpy as np

ORACLE_DIR = pathlib.Path(__file__).resolve().parent
LIB_PATH = ORACLE_DIR / "_build" / "libnbg_oracle.so"

_lib: ctypes.CDLL | None = None


def build() -> pathlib.Path:
    """Compile the C oracle (make -C oracle)."""
    proc = subprocess.run(["make", "-C", str(ORACLE_DIR)], capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{proc.stdout}\n{proc.stderr}")
    return LIB_PATH


def use_native_build(out_dir: str | pathlib.Path) -> pathlib.Path:
    """Compile the C oracle for the host it runs on (-O3 -march=native
    -ffp-contract=off, OpenMP) into ``out_dir`` and make ``restated`` use it.

    For the CPU baseline leg of bench.py: the portable build under _build/
    serves the tests, this one is tuned for the machine that times it.
    """
    global _lib
    out = pathlib.Path(out_dir) / "libnbg_oracle_native.so"
    out.parent.mkdir(parents=True, exist_ok=True)
    cmd = ["gcc", "-O3", "-march=native", "-ffp-contract=off", "-fno-fast-math", "-fopenmp", "-fPIC",
           "-shared", "-o", str(out), str(ORACLE_DIR / "nbg_oracle.c"), "-lm"]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"native oracle build failed:\n{proc.stderr}")
    _lib = ctypes.CDLL(str(out))
    return out


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists() or LIB_PATH.stat().st_mtime < (ORACLE_DIR / "nbg_oracle.c").stat().st_mtime:
            build()
        _lib = ctypes.CDLL(str(LIB_PATH))
    return _lib


_SFX = {np.dtype(np.float32): "f32", np.dtype(np.float64): "f64", np.dtype(np.int32): "i32", np.dtype(np.int64): "i64"}


def _p(a: np.ndarray) -> ctypes.c_void_p:
    return ctypes.c_void_p(a.ctypes.data)


def _c2(a: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(a)
    if a.ndim == 1:
        a = a[None, :]
    assert a.ndim == 2
    return a


I64 = ctypes.c_int64


class restated:
    """The reference kernels restated in C (same types and operation order)."""

    @staticmethod
    def move(func: str, x: np.ndarray, window: int, min_count: int | None = None, y: np.ndarray | None = None):
        x = _c2(x)
        mc = window if min_count is None else min_count
        out = np.empty_like(x)
        rows, n = x.shape
        sfx = _SFX[x.dtype]
        fn = getattr(lib(), f"orc_{func}_{sfx}")
        if y is None:
            fn(_p(x), _p(out), I64(rows), I64(n), I64(n), I64(n), I64(window), I64(mc))
        else:
            y = _c2(y).astype(x.dtype, copy=False)
            fn(_p(x), _p(y), _p(out), I64(rows), I64(n), I64(n), I64(n), I64(n), I64(window), I64(mc))
        return out

    @staticmethod
    def move_exp(func: str, x: np.ndarray, alpha, min_weight: float = 0.0, y: np.ndarray | None = None):
        """alpha: scalar, (n,) row shared by all rows, or (rows, n); in x's dtype."""
        x = _c2(x)
        rows, n = x.shape
        al = np.asarray(alpha, dtype=x.dtype)
        if al.ndim == 0:
            al = np.full((1, n), al, dtype=x.dtype)
            lda = 0
        elif al.ndim == 1:
            al = al[None, :]
            lda = 0
        else:
            lda = n
        al = np.ascontiguousarray(al)
        out = np.empty_like(x)
        sfx = _SFX[x.dtype]
        fn = getattr(lib(), f"orc_{func}_{sfx}")
        mw = ctypes.c_float(min_weight) if x.dtype == np.float32 else ctypes.c_double(min_weight)
        if y is None:
            fn(_p(x), _p(al), _p(out), I64(rows), I64(n), I64(n), I64(lda), I64(n), mw)
        else:
            y = _c2(y).astype(x.dtype, copy=False)
            fn(_p(x), _p(y), _p(al), _p(out), I64(rows), I64(n), I64(n), I64(n), I64(lda), I64(n), mw)
        return out

    @staticmethod
    def fill(func: str, x: np.ndarray, limit: int | None = None):
        x = _c2(x)
        rows, n = x.shape
        lim = n if limit is None else limit
        out = np.empty_like(x)
        getattr(lib(), f"orc_{func}_{_SFX[x.dtype]}")(_p(x), _p(out), I64(rows), I64(n), I64(n), I64(n), I64(lim))
        return out

    @staticmethod
    def group(func: str, values: np.ndarray, labels: np.ndarray, num_labels: int, ddof: int | None = None):
        v = _c2(values)
        lab = np.ascontiguousarray(labels, dtype=np.int64).ravel()
        rows, n = v.shape
        assert lab.shape == (n,)
        out = np.empty((rows, num_labels), dtype=v.dtype)
        fn = getattr(lib(), f"orc_{func}_{_SFX[v.dtype]}")
        args = [_p(v), _p(lab), _p(out), I64(rows), I64(n), I64(n), I64(num_labels)]
        if ddof is not None:
            args.append(I64(ddof))
        fn(*args)
        return out

    @staticmethod
    def reduce(func: str, x: np.ndarray, ddof: int = 1) -> np.ndarray:
        """Plain reductions over the last axis (numbagg/funcs.py:23-228); the
        nanargmax/nanargmin rows without a valid value give -1."""
        v = _c2(x)
        rows, n = v.shape
        if func in ("allnan", "anynan"):
            odt = np.uint8
        elif func in ("nancount", "nanargmax", "nanargmin"):
            odt = np.int64
        elif func in ("nanmax", "nanmin") and v.dtype.kind == "i":
            odt = np.int64
        else:
            odt = v.dtype
        out = np.empty(rows, dtype=odt)
        fn = getattr(lib(), f"orc_{func}_{_SFX[v.dtype]}")
        args = [_p(v), _p(out), I64(rows), I64(n), I64(n)]
        if func in ("nanvar", "nanstd"):
            args.append(I64(ddof))
        fn(*args)
        return out.astype(bool) if odt == np.uint8 else out


    @staticmethod
    def matrix(func: str, x: np.ndarray, window: int | None = None, min_count: int | None = None, alpha=None,
               min_weight: float = 0.0) -> np.ndarray:
        """Matrix functions on a 3-D batch: static (batch, vars, obs) ->
        (batch, vars, vars); moving and exponential (batch, obs, vars) ->
        (batch, obs, vars, vars). alpha: scalar, (obs,) or (batch, obs)."""
        x = np.ascontiguousarray(x)
        if x.ndim == 2:
            x = x[None]
        sfx = _SFX[x.dtype]
        fn = getattr(lib(), f"orc_{func}_{sfx}")
        bsz = x.shape[0]
        if func in ("nancorrmatrix", "nancovmatrix"):
            nv, nobs = x.shape[1:]
            out = np.empty((bsz, nv, nv), dtype=x.dtype)
            fn(_p(x), _p(out), I64(bsz), I64(nv), I64(nobs))
            return out
        nobs, nv = x.shape[1:]
        out = np.empty((bsz, nobs, nv, nv), dtype=x.dtype)
        if func in ("move_corrmatrix", "move_covmatrix"):
            mc = window if min_count is None else min_count
            fn(_p(x), _p(out), I64(bsz), I64(nobs), I64(nv), I64(window), I64(mc))
            return out
        al = np.asarray(alpha, dtype=x.dtype)
        if al.ndim == 0:
            al, lda = np.full((1, nobs), al, dtype=x.dtype), 0
        elif al.ndim == 1:
            al, lda = al[None, :], 0
        else:
            lda = nobs
        al = np.ascontiguousarray(al)
        mw = ctypes.c_float(min_weight) if x.dtype == np.float32 else ctypes.c_double(min_weight)
        fn(_p(x), _p(al), I64(lda), _p(out), I64(bsz), I64(nobs), I64(nv), mw)
        return out


_MOVE_KIND = {"move_sum": 0, "move_mean": 1, "move_var": 2, "move_std": 3, "move_cov": 4, "move_corr": 5}
_EXP_KIND = {
    "move_exp_nancount": 0,
    "move_exp_nanmean": 1,
    "move_exp_nansum": 2,
    "move_exp_nanvar": 3,
    "move_exp_nanstd": 4,
    "move_exp_nancov": 5,
    "move_exp_nancorr": 6,
}


class exact:
    """Extended-precision oracles; results are float64."""

    @staticmethod
    def move(func: str, x: np.ndarray, window: int, min_count: int | None = None, y: np.ndarray | None = None):
        x = _c2(x).astype(np.float64)
        rows, n = x.shape
        mc = window if min_count is None else min_count
        out = np.empty_like(x)
        yp = ctypes.c_void_p(0)
        if y is not None:
            y = _c2(y).astype(np.float64)
            yp = _p(y)
        lib().orc_exact_move(
            ctypes.c_int(_MOVE_KIND[func]), _p(x), yp, _p(out), I64(rows), I64(n), I64(window), I64(mc)
        )
        return out

    @staticmethod
    def move_exp(func: str, x: np.ndarray, alpha, min_weight: float = 0.0, y: np.ndarray | None = None, loop_dtype=None):
        """alpha/min_weight are rounded to ``loop_dtype`` (x's dtype by default) first."""
        x0 = _c2(x)
        loop = np.dtype(loop_dtype or x0.dtype)
        x = x0.astype(np.float64)
        rows, n = x.shape
        al = np.asarray(np.asarray(alpha, dtype=loop), dtype=np.float64)
        if al.ndim == 0:
            al, lda = np.full((1, n), al), 0
        elif al.ndim == 1:
            al, lda = al[None, :], 0
        else:
            lda = n
        al = np.ascontiguousarray(al)
        mw = float(np.asarray(min_weight, dtype=loop))
        out = np.empty_like(x)
        yp = ctypes.c_void_p(0)
        if y is not None:
            y = _c2(y).astype(np.float64)
            yp = _p(y)
        lib().orc_exact_move_exp(
            ctypes.c_int(_EXP_KIND[func]), _p(x), yp, _p(al), I64(lda), _p(out), I64(rows), I64(n),
            ctypes.c_double(mw),
        )
        return out

    @staticmethod
    def fill(func: str, x: np.ndarray, limit: int | None = None) -> np.ndarray:
        """Closed form: out[i] = x[j], j the last valid index at or before i in
        scan order, when i - j <= limit; canonical NaN otherwise."""
        x = _c2(x)
        rows, n = x.shape
        lim = n if limit is None else limit
        xs = x[:, ::-1] if func == "bfill" else x
        idx = np.arange(n)
        valid = ~np.isnan(xs)
        last = np.where(valid, idx[None, :], -1)
        last = np.maximum.accumulate(last, axis=1)
        ok = (last >= 0) & ((idx[None, :] - last) <= lim)
        out = np.full_like(xs, np.nan)
        rr = np.repeat(np.arange(rows)[:, None], n, axis=1)
        out[ok] = xs[rr[ok], last[ok]]
        return out[:, ::-1].copy() if func == "bfill" else out

    @staticmethod
    def group(func: str, values: np.ndarray, labels: np.ndarray, num_labels: int, ddof: int = 1) -> np.ndarray:
        """float64 closed forms (two-pass variance) per (row, label); products are
        float64 products in index order, min/max the value at the first index of
        the extreme."""
        v = _c2(values)
        lab = np.asarray(labels, dtype=np.int64).ravel()
        rows, n = v.shape
        out = np.full((rows, num_labels), np.nan)
        keep = (lab >= 0) & (lab < num_labels)
        for r in range(rows):
            vr = v[r].astype(np.float64)
            ok = keep & ~np.isnan(vr)
            lr, xr = lab[ok], vr[ok]
            cnt = np.bincount(lr, minlength=num_labels).astype(np.float64)
            if func == "group_nancount":
                out[r] = cnt
                continue
            s = np.bincount(lr, weights=xr, minlength=num_labels)
            if func == "group_nansum":
                out[r] = s
            elif func == "group_nanmean":
                with np.errstate(invalid="ignore", divide="ignore"):
                    out[r] = np.where(cnt > 0, s / cnt, np.nan)
            elif func in ("group_nanvar", "group_nanstd"):
                with np.errstate(invalid="ignore", divide="ignore"):
                    m = s / np.where(cnt > 0, cnt, 1)
                    m2 = np.bincount(lr, weights=(xr - m[lr]) ** 2, minlength=num_labels)
                    den = cnt - ddof
                    # an empty label is NaN whatever ddof: the reference's
                    # (0 - 0**2/0) / (0 - ddof) (numbagg/grouped.py:166-174)
                    var = np.where((den > 0) & (cnt > 0), m2 / np.where(den > 0, den, 1), np.nan)
                out[r] = np.sqrt(var) if func == "group_nanstd" else var
            elif func == "group_nansum_of_squares":
                out[r] = np.bincount(lr, weights=xr * xr, minlength=num_labels)
            elif func in ("group_nanany", "group_nanall"):
                hit = (xr != 0) if func == "group_nanany" else (xr == 0)
                seen = np.bincount(lr[hit], minlength=num_labels) > 0
                out[r] = seen if func == "group_nanany" else ~seen
            elif func == "group_nanprod":
                out[r] = 1.0
                order = np.argsort(lr, kind="stable")
                bounds = np.searchsorted(lr[order], np.arange(num_labels + 1))
                for k in range(num_labels):
                    out[r, k] = np.prod(xr[order[bounds[k] : bounds[k + 1]]])
            elif func in ("group_nanargmax", "group_nanargmin", "group_nanfirst", "group_nanlast",
                          "group_nanmin", "group_nanmax"):
                pos = np.nonzero(ok)[0]
                for k in range(num_labels):
                    sel = pos[lr == k]
                    if sel.size == 0:
                        continue
                    vals = vr[sel]
                    # np.argmax/argmin return the first of equal extremes (-0.0 == 0.0)
                    if func == "group_nanargmax":
                        out[r, k] = sel[np.argmax(vals)]
                    elif func == "group_nanmax":
                        out[r, k] = vals[np.argmax(vals)]
                    elif func == "group_nanmin":
                        out[r, k] = vals[np.argmin(vals)]
                    elif func == "group_nanargmin":
                        out[r, k] = sel[np.argmin(vals)]
                    elif func == "group_nanfirst":
                        out[r, k] = vr[sel[0]]
                    else:
                        out[r, k] = vr[sel[-1]]
            else:
                raise ValueError(func)
        return out

    @staticmethod
    def reduce(func: str, x: np.ndarray, ddof: int = 1) -> np.ndarray:
        """float64 closed forms over the last axis: exactly rounded sums
        (math.fsum), two-pass variance about the fsum mean."""
        v = _c2(x).astype(np.float64)
        out = np.full(v.shape[0], np.nan)
        for r, row in enumerate(v):
            ok = row[~np.isnan(row)]
            if func == "nansum":
                out[r] = math.fsum(ok)
            elif func == "nanmean":
                out[r] = math.fsum(ok) / ok.size if ok.size else np.nan
            elif func in ("nanvar", "nanstd"):
                if ok.size > ddof:
                    m = math.fsum(ok) / ok.size
                    var = math.fsum((ok - m) ** 2) / (ok.size - ddof)
                    out[r] = np.sqrt(var) if func == "nanstd" else var
            else:
                raise ValueError(func)
        return out

    @staticmethod
    def matrix(func: str, x: np.ndarray, window: int | None = None, min_count: int | None = None, alpha=None,
               min_weight: float = 0.0, loop_dtype=None) -> np.ndarray:
        """Matrix functions pair by pair (float64 results): static ones by a
        two-pass long-double evaluation over each pair's complete observations;
        moving and exponential ones by the two-series oracles above applied to
        every (i, j) pair (a pair's statistics are those of move_cov/move_corr
        and move_exp_nancov/nancorr on columns i and j)."""
        x0 = np.ascontiguousarray(x)
        if x0.ndim == 2:
            x0 = x0[None]
        loop = np.dtype(loop_dtype or x0.dtype)
        xf = x0.astype(np.float64)
        bsz = xf.shape[0]
        if func in ("nancorrmatrix", "nancovmatrix"):
            nv, nobs = xf.shape[1:]
            out = np.empty((bsz, nv, nv))
            lib().orc_exact_matrix(ctypes.c_int(1 if func == "nancorrmatrix" else 0), _p(xf), _p(out), I64(bsz),
                                   I64(nv), I64(nobs))
            return out
        nobs, nv = xf.shape[1:]
        cols = np.moveaxis(xf, 2, 1)  # (batch, vars, obs)
        xi = np.repeat(cols[:, :, None, :], nv, axis=2).reshape(-1, nobs)
        xj = np.repeat(cols[:, None, :, :], nv, axis=1).reshape(-1, nobs)
        if func in ("move_corrmatrix", "move_covmatrix"):
            mc = window if min_count is None else min_count
            if func == "move_corrmatrix":
                mc = max(mc, 2)
            res = exact.move("move_corr" if func == "move_corrmatrix" else "move_cov", xi, window, mc, y=xj)
        else:
            al = np.asarray(alpha, dtype=loop)
            if al.ndim == 2:
                al = np.repeat(al, nv * nv, axis=0)
            res = exact.move_exp("move_exp_nancorr" if func == "move_exp_nancorrmatrix" else "move_exp_nancov",
                                 xi, al, min_weight, y=xj, loop_dtype=loop)
        return np.moveaxis(res.reshape(bsz, nv, nv, nobs), 3, 1).copy()

    @staticmethod
    def factorize(values: np.ndarray, sort: bool = True):
        """Closed form of pandas.factorize for 1-D labels: np.unique over the
        non-NaN values (-0.0 folded onto +0.0), the first occurrence of each
        value as its representative; codes -1 for NaN; uniques ascending or in
        order of first appearance."""
        v = np.asarray(values).ravel()
        nan = np.isnan(v) if v.dtype.kind == "f" else np.zeros(v.shape, bool)
        pos = np.nonzero(~nan)[0]
        keys = v[pos] + 0 if v.dtype.kind == "f" else v[pos]
        uk, first, inv = np.unique(keys, return_index=True, return_inverse=True)
        if not sort:
            order = np.argsort(first, kind="stable")
            rank = np.empty_like(order)
            rank[order] = np.arange(order.size)
            inv, first = rank[inv], first[order]
        codes = np.full(v.shape, -1, dtype=np.int64)
        codes[pos] = inv
        return codes, v[pos[first]]
