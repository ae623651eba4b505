"""Host-side operator classes: numbagg's decorator semantics over HIP kernels.

Each class restates the argument handling of one numbagg decorator
(numbagg/decorators.py) — defaults, validation, error types and messages, axis
handling, dtype selection — and then, instead of dispatching to a numba gufunc,
normalises the operands to a (rows, n) device view with a contiguous core axis
and calls the matching ``nbg_*`` entry point of the C ABI.

The decorated Python functions carry only the public name and docstring; their
bodies are never executed. There is no CPU fallback.
"""

from __future__ import annotations

import ctypes
import functools
import inspect
import logging
import os
from collections.abc import Callable
from typing import Any

import numpy as np
import torch
from numpy.lib.array_utils import normalize_axis_tuple

from . import _device as dev
from . import _lib
from . import plain_tuning

logger = logging.getLogger(__name__)


def _loop_dtype(dtypes: list[np.dtype], func_name: str) -> np.dtype:
    """The float loop NumPy's gufunc type resolution would pick.

    numbagg registers a float32 loop then a float64 loop (e.g. moving.py:12-14);
    NumPy takes the first loop every operand can be cast to safely.
    """
    for loop in (np.dtype(np.float32), np.dtype(np.float64)):
        if all(np.can_cast(dt, loop, casting="safe") for dt in dtypes):
            return loop
    raise TypeError(
        f"ufunc '{func_name}' not supported for the input types "
        f"{tuple(str(d) for d in dtypes)}, and the inputs could not be safely "
        "coerced to any supported types according to the casting rule ''safe''"
    )


def _shape(a) -> tuple[int, ...]:
    if isinstance(a, torch.Tensor):
        return tuple(a.shape)
    return np.shape(a)


def _is_int_scalar(v) -> bool:
    return isinstance(v, (int, np.integer)) and not isinstance(v, (bool, np.bool_))


class _Op:
    """Common base: name/doc copying and the ``numbagg.<name>`` repr."""

    def __init__(self, func: Callable[..., Any]) -> None:
        self.func = func
        functools.wraps(func)(self)
        self.__signature__ = inspect.signature(self.__call__)
        self.entry = "nbg_" + func.__name__

    def __repr__(self) -> str:
        return f"numbagg.{self.__name__}"  # type: ignore[attr-defined]

    @classmethod
    def wrap(cls, *args, **kwargs):
        """Decorate a documentation stub (numbagg/decorators.py:122-127)."""
        return lambda func: cls(func, *args, **kwargs)

    @staticmethod
    def _no_extra(kwargs: dict) -> None:
        if kwargs:
            raise TypeError(f"unsupported keyword arguments: {sorted(kwargs)}")

    @classmethod
    def _take_out(cls, kwargs: dict):
        """The gufunc ``out=`` argument the reference forwards with **kwargs
        (numbagg/decorators.py:341, :414, :487, :731, :807, :1089): an array
        (or a 1-tuple holding one) that receives the result."""
        dest = kwargs.pop("out", None)
        cls._no_extra(kwargs)
        if isinstance(dest, tuple):
            if len(dest) != 1:
                raise ValueError("The 'out' tuple must have exactly one entry")
            dest = dest[0]
        if dest is not None and not isinstance(dest, (np.ndarray, torch.Tensor)):
            raise TypeError("return arrays must be of ArrayType")
        return dest

    def _deliver(self, res, dest):
        """Write ``res`` into ``dest`` as NumPy's gufunc output does (exact
        shape, casting 'same_kind') and return ``dest``."""
        if dest is None:
            return res
        if isinstance(dest, torch.Tensor):
            r = res if isinstance(res, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(res))
            if tuple(dest.shape) != tuple(r.shape):
                raise ValueError(f"output operand with shape {tuple(dest.shape)} does not match the result shape "
                                 f"{tuple(r.shape)}")
            dest.copy_(r)
            return dest
        r = res.cpu().numpy() if isinstance(res, torch.Tensor) else np.asarray(res)
        if dest.shape != r.shape:
            raise ValueError(f"non-broadcastable output operand with shape {dest.shape} doesn't match the "
                             f"broadcast shape {r.shape}")
        if not np.can_cast(r.dtype, dest.dtype, casting="same_kind"):
            raise TypeError(f"Cannot cast ufunc '{self.__name__}' output from {r.dtype!r} to {dest.dtype!r} with "
                            "casting rule 'same_kind'")
        np.copyto(dest, r, casting="same_kind")
        return dest


def _single_axis(axis, arr: tuple, func, capital: bool):
    """Axis-tuple rules shared by ndmove/ndmoveexp (decorators.py:318-332, :385-396)."""
    if isinstance(axis, tuple):
        if axis == ():
            if len(arr) > 1:
                raise ValueError(
                    "`axis` cannot be an empty tuple when passing more than one array; "
                    "since we default to returning the input."
                )
            return None
        if len(axis) > 1:
            word = "Only" if capital else "only"
            raise ValueError(f"{word} one axis can be passed to {func}; got {axis}")
        (axis,) = axis
    return axis


def _broadcast_rows(tensors: list[torch.Tensor], axes: list[int]):
    """Move each operand's core axis last and broadcast the outer dimensions.

    Returns (list of (rows, n) views with ld, outer_shape, n).
    """
    moved = [t.movedim(ax, -1) for t, ax in zip(tensors, axes)]
    n = moved[0].shape[-1]
    for m in moved[1:]:
        if m.shape[-1] != n:
            raise ValueError(
                f"core dimension mismatch: operands have core lengths {[x.shape[-1] for x in moved]}"
            )
    outer = torch.broadcast_shapes(*[tuple(m.shape[:-1]) for m in moved])
    views = []
    for m in moved:
        if tuple(m.shape[:-1]) != tuple(outer):
            m = m.expand(*outer, n)
        views.append(dev.rows_view(m))
    return views, tuple(outer), n


class ndmove(_Op):
    """Moving-window operator (numbagg/decorators.py:275-341).

    Calls ``nbg_<name>(dtype, inputs..., out, rows, n, ld_inputs..., ld_out,
    window, min_count, stream)``.
    """

    def __init__(self, func: Callable[..., Any], nargs: int = 1) -> None:
        super().__init__(func)
        self.nargs = nargs

    def __call__(
        self,
        *arr,
        window: int,
        min_count: int | None = None,
        axis: int | tuple[int, ...] = -1,
        **kwargs,
    ):
        dest = self._take_out(kwargs)
        if min_count is None:
            min_count = window
        elif min_count < 0:
            raise ValueError(f"min_count must be positive: {min_count}")
        axis = _single_axis(axis, arr, self.func, capital=False)
        if axis is None:
            return arr[0]
        if len(arr) != self.nargs:
            raise TypeError(f"{self.__name__} takes {self.nargs} array(s), got {len(arr)}")
        if not 0 < window <= _shape(arr[0])[axis]:
            raise ValueError(f"window not in valid range: {window}")
        if not _is_int_scalar(window) or not _is_int_scalar(min_count):
            raise TypeError(
                f"ufunc '{self.__name__}' not supported for the input types: window and "
                "min_count must be integers"
            )
        as_tensor = any(dev.is_tensor(a) for a in arr)
        loop = _loop_dtype([dev.np_dtype(a) for a in arr], self.__name__)
        tensors = [dev.to_device(a, loop) for a in arr]
        views, outer, n = _broadcast_rows(tensors, [axis] * len(tensors))
        rows = views[0][1]
        out = torch.empty((rows, n), dtype=dev.torch_dtype(loop), device=tensors[0].device)
        if rows and n:
            args = [ctypes.c_int(dev.nbg_code(out))]
            args += [dev.ptr(v[0]) for v in views]
            args += [dev.ptr(out), rows, n]
            args += [v[3] for v in views]
            args += [n, int(window), int(min_count), dev.stream_handle()]
            with plain_tuning.applied(self.__name__, loop, rows, n, int(window)):
                _lib.call(self.entry, *args)
        # NumPy places the output's core dimension at `axis` as well.
        res = out.view(*outer, n).movedim(-1, axis)
        return self._deliver(dev.to_output(res, as_tensor), dest)


class ndmoveexp(_Op):
    """Exponentially weighted moving operator (numbagg/decorators.py:344-414).

    ``alpha`` may be a scalar, a 1-D array along the moving axis (shared by all
    rows) or an N-D array laid out like the input. Calls
    ``nbg_<name>(x_dtype, dtype, inputs..., alpha, ld_alpha, alpha_scalar, out,
    rows, n, ld_inputs..., ld_out, min_weight, ws, ws_bytes, stream)``.
    """

    def __init__(self, func: Callable[..., Any], nargs: int = 1) -> None:
        super().__init__(func)
        self.nargs = nargs

    def __call__(
        self,
        *arr,
        alpha,
        min_weight: float = 0,
        axis: int | tuple[int, ...] = -1,
        **kwargs,
    ):
        dest = self._take_out(kwargs)
        axis = _single_axis(axis, arr, self.func, capital=True)
        if axis is None:
            return arr[0]
        if len(arr) != self.nargs:
            raise TypeError(f"{self.__name__} takes {self.nargs} array(s), got {len(arr)}")
        n_axis = _shape(arr[0])[axis]
        as_tensor = any(dev.is_tensor(a) for a in arr)

        alpha_is_array = isinstance(alpha, (np.ndarray, torch.Tensor))
        if not alpha_is_array:
            # np.broadcast_to(alpha, n) in the reference: the dtype of the scalar
            # (a Python float is float64) takes part in the loop selection.
            alpha_dtype = np.asarray(alpha).dtype
            alpha_mode = "scalar"
        else:
            alpha_dtype = dev.np_dtype(alpha)
            alpha_mode = "row" if alpha.ndim == 1 else "full"
        dtypes = [dev.np_dtype(a) for a in arr] + [alpha_dtype]
        if isinstance(min_weight, np.generic):
            dtypes.append(min_weight.dtype)
        loop = _loop_dtype(dtypes, self.__name__)
        mw = float(np.asarray(min_weight, dtype=loop))

        in_dtypes = {dev.np_dtype(a) for a in arr}
        # float32 data on the float64 loop is widened inside the kernel.
        fused = loop == np.float64 and in_dtypes == {np.dtype(np.float32)}
        x_dtype = np.dtype(np.float32) if fused else loop
        tensors = [dev.to_device(a, x_dtype) for a in arr]
        views, outer, n = _broadcast_rows(tensors, [axis] * len(tensors))
        rows = views[0][1]
        device = tensors[0].device

        alpha_ptr, ld_alpha, alpha_scalar, alpha_keep = ctypes.c_void_p(0), 0, 0.0, None
        if alpha_mode == "scalar":
            alpha_scalar = float(np.asarray(alpha, dtype=loop))
        elif alpha_mode == "row":
            if alpha.shape[0] != n_axis:
                raise ValueError(
                    f"{self.__name__}: alpha has length {alpha.shape[0]} but the core "
                    f"dimension has length {n_axis}"
                )
            alpha_keep = dev.to_device(alpha, loop).contiguous()
            alpha_ptr, ld_alpha = dev.ptr(alpha_keep), 0
        else:
            a_t = dev.to_device(alpha, loop).movedim(axis, -1)
            if a_t.shape[-1] != n:
                raise ValueError(f"{self.__name__}: alpha core dimension mismatch")
            alpha_keep = a_t.expand(*outer, n).contiguous().view(max(rows, 1), n)
            alpha_ptr, ld_alpha = dev.ptr(alpha_keep), n

        out = torch.empty((rows, n), dtype=dev.torch_dtype(loop), device=device)
        if rows and n:
            code_in = _lib.NBG_F32 if x_dtype == np.float32 else _lib.NBG_F64
            ws_bytes = _lib.query("nbg_move_exp_workspace_bytes", rows, n, code_in)
            ws, ws_bytes = dev.workspace(ws_bytes, device)
            args = [ctypes.c_int(code_in), ctypes.c_int(dev.nbg_code(out))]
            args += [dev.ptr(v[0]) for v in views]
            args += [alpha_ptr, ld_alpha, ctypes.c_double(alpha_scalar), dev.ptr(out), rows, n]
            args += [v[3] for v in views]
            args += [n, ctypes.c_double(mw), dev.ptr(ws), ws_bytes, dev.stream_handle()]
            with plain_tuning.applied(self.__name__, loop, rows, n):
                _lib.call(self.entry, *args)
        del alpha_keep
        # NumPy places the output's core dimension at `axis` as well.
        res = out.view(*outer, n).movedim(-1, axis)
        return self._deliver(dev.to_output(res, as_tensor), dest)


class ndfill(_Op):
    """Forward/backward fill operator (numbagg/decorators.py:417-487).

    Float rows run the HIP kernel; integer rows have no NaN, so the reference
    copies them through unchanged, and so does this class (on the device).
    Calls ``nbg_<name>(dtype, x, out, rows, n, ld_x, ld_out, limit, ws,
    ws_bytes, stream)``.
    """

    def __call__(self, arr, *, limit: None | int = None, axis: int = -1, **kwargs):
        dest = self._take_out(kwargs)
        shape = _shape(arr)
        if limit is None:
            limit = shape[axis]
        if limit < 0:
            raise ValueError(f"`limit` must be positive: {limit}")
        dt = dev.np_dtype(arr)
        if not np.issubdtype(dt, np.number):
            raise TypeError(f"Unsupported dtype for fill operation: {dt}")
        as_tensor = dev.is_tensor(arr)
        if not np.issubdtype(dt, np.floating):
            t = dev.to_device(arr)
            return self._deliver(dev.to_output(t.clone(), as_tensor), dest)
        # float16 rows are filled through float32, which round-trips exactly.
        work = np.dtype(np.float32) if dt == np.float16 else dt
        t = dev.to_device(arr, work)
        moved = t.movedim(axis, -1)
        v2, rows, n, ld = dev.rows_view(moved)
        out = torch.empty((rows, n), dtype=t.dtype, device=t.device)
        if rows and n:
            # the workspace query and the call see the same knob settings
            with plain_tuning.applied(self.__name__, work, rows, n):
                ws_bytes = _lib.query("nbg_fill_workspace_bytes", rows, n, dev.nbg_code(t))
                ws, ws_bytes = dev.workspace(ws_bytes, t.device)
                _lib.call(
                    self.entry,
                    ctypes.c_int(dev.nbg_code(t)),
                    dev.ptr(v2),
                    dev.ptr(out),
                    rows,
                    n,
                    ld,
                    n,
                    int(limit),
                    dev.ptr(ws),
                    ws_bytes,
                    dev.stream_handle(),
                )
        res = out.view(moved.shape).movedim(-1, axis)
        if work != dt:
            res = res.to(dev.torch_dtype(dt))
        return self._deliver(dev.to_output(res, as_tensor), dest)


# Group-by strategy codes of the C ABI (nbg_group_*).
GROUP_STRATEGY_AUTO = 0
GROUP_STRATEGY_LDS = 1
GROUP_STRATEGY_GLOBAL = 2


def _group_strategy() -> int:
    """Strategy override from NBG_GROUP_STRATEGY ("lds" / "global"), else auto."""
    v = os.environ.get("NBG_GROUP_STRATEGY", "").lower()
    return {"lds": GROUP_STRATEGY_LDS, "global": GROUP_STRATEGY_GLOBAL}.get(v, GROUP_STRATEGY_AUTO)


class groupndreduce(_Op):
    """Grouped reduction operator (numbagg/decorators.py:490-674).

    Supports the reference's three axis modes (None, int, tuple). Values are
    reduced into ``(*broadcast_shape, num_labels)`` by
    ``nbg_<name>(vdtype, values, ldtype, labels, out, rows, n, ld_v, num_labels,
    [ddof,] strategy, ws, ws_bytes, stream)``. Labels < 0 are skipped; labels
    >= num_labels (undefined behaviour in the reference) are skipped as well.
    """

    def __init__(
        self,
        func: Callable[..., Any],
        *,
        supports_ddof: bool = False,
        supports_bool: bool = True,
        supports_ints: bool = True,
    ) -> None:
        self.supports_bool = supports_bool
        self.supports_ints = supports_ints
        self.supports_ddof = supports_ddof
        super().__init__(func)

    def __call__(
        self,
        values,
        labels,
        *,
        ddof: int = 1,
        num_labels: int | None = None,
        axis: int | tuple[int, ...] | None = None,
    ):
        as_tensor = dev.is_tensor(values)
        if not dev.is_tensor(values):
            values = np.asarray(values)
        if not dev.is_tensor(labels):
            labels = np.asarray(labels)
        ldt = dev.np_dtype(labels)
        if ldt.kind not in "i":
            raise TypeError(
                "labels must be an integer array; it's expected to have already been "
                "factorized with a function such as `pd.factorize`"
            )
        vdt = dev.np_dtype(values)
        vshape = tuple(values.shape)
        lshape = tuple(labels.shape)
        if vdt == np.bool_:
            if not self.supports_bool:
                func_name = getattr(self.func, "__name__", "<unknown>")
                raise TypeError(
                    f"{func_name} does not support boolean input. Convert to a numeric type first."
                )
            vdt = np.dtype(np.int32)
        if num_labels is None:
            if dev.is_tensor(labels):
                num_labels = int(labels.max().item()) + 1
            else:
                num_labels = int(np.max(labels)) + 1
        if np.issubdtype(vdt, np.integer) and not self.supports_ints:
            vdt = np.dtype(np.float64)
        result_dtype = vdt

        ndim = len(vshape)
        if axis is None:
            if vshape != lshape:
                raise ValueError(
                    "axis required if values and labels have different "
                    f"shapes: {vshape} vs {lshape}"
                )
            core_axes = list(range(ndim))
        elif isinstance(axis, (int, np.integer)):
            ax = int(axis)
            if lshape != (vshape[ax],):
                raise ValueError(
                    "values must have same shape along axis as labels: "
                    f"{(vshape[ax],)} vs {lshape}"
                )
            core_axes = [ax % ndim]
        else:
            values_shape = tuple(vshape[ax] for ax in axis)
            if lshape != values_shape:
                raise ValueError(
                    f"values must have same shape along axis as labels: {values_shape} vs {lshape}"
                )
            core_axes = [ax % ndim for ax in axis]
        broadcast_shape = tuple(s for i, s in enumerate(vshape) if i not in core_axes)

        kdt = self._kernel_value_dtype(result_dtype)
        v = dev.to_device(values, kdt)
        if v.dtype == torch.bool:
            v = v.to(torch.int32)
        v = v.movedim(core_axes, list(range(ndim - len(core_axes), ndim)))
        n = int(np.prod([vshape[a] for a in core_axes], dtype=np.int64))
        rows = int(np.prod(broadcast_shape, dtype=np.int64))
        v2 = v.reshape(rows, n)
        if v2.stride(-1) != 1 or (rows > 1 and v2.stride(0) < n):
            v2 = v2.contiguous()
        ld_v = int(v2.stride(0)) if rows > 1 else n
        lab = dev.to_device(labels)
        lab = lab.to(torch.int32 if lab.dtype in (torch.int8, torch.int16) else lab.dtype)
        lab = lab.reshape(n).contiguous()

        out = torch.empty((rows, num_labels), dtype=dev.torch_dtype(kdt), device=v2.device)
        if rows and num_labels:
            strategy = _group_strategy()
            ws_bytes = _lib.query(
                "nbg_group_workspace_bytes", dev.nbg_code(v2), rows, n, num_labels, strategy
            )
            ws, ws_bytes = dev.workspace(ws_bytes, v2.device)
            args = [
                ctypes.c_int(dev.nbg_code(v2)),
                dev.ptr(v2),
                ctypes.c_int(dev.nbg_code(lab)),
                dev.ptr(lab),
                dev.ptr(out),
                rows,
                n,
                ld_v,
                int(num_labels),
            ]
            if self.supports_ddof:
                args.append(int(ddof))
            args += [ctypes.c_int(strategy), dev.ptr(ws), ws_bytes, dev.stream_handle()]
            with plain_tuning.applied(self.__name__, kdt, rows, n, int(num_labels)):
                _lib.call(self.entry, *args)
        res = out.view(*broadcast_shape, num_labels)
        if kdt != result_dtype:
            res = res.to(dev.torch_dtype(result_dtype))
        return dev.to_output(res, as_tensor)

    @staticmethod
    def _kernel_value_dtype(dt: np.dtype) -> np.dtype:
        """Value dtype the kernels run: float32/float64/int32/int64.

        Narrow integers are widened to int32 (wrap-around sums and index outputs
        are narrowed back afterwards, which reproduces the reference's modular
        integer arithmetic); float16 runs in float32.
        """
        dt = np.dtype(dt)
        if dt in (np.dtype(np.float32), np.dtype(np.float64), np.dtype(np.int32), np.dtype(np.int64)):
            return dt
        if dt == np.float16:
            return np.dtype(np.float32)
        if dt.kind in "iub" and dt.itemsize <= 4 and dt != np.uint32:
            return np.dtype(np.int32)
        if dt == np.uint32:
            return np.dtype(np.int64)
        raise TypeError(f"numbagg_amd: unsupported values dtype {dt}")


def _first_safe_loop(dt: np.dtype, loops: list[np.dtype], func_name: str) -> np.dtype:
    """NumPy's loop choice for a gufunc with the given loops: the first one the
    input casts to safely (what numba's guvectorize registers, in order)."""
    for loop in loops:
        if np.can_cast(dt, loop, casting="safe"):
            return loop
    raise TypeError(
        f"ufunc '{func_name}' not supported for the input types ({dt},), and the inputs "
        "could not be safely coerced to any supported types according to the casting rule ''safe''"
    )


def _reduce_rows(t: torch.Tensor, axes: tuple[int, ...]):
    """Move `axes` (in that order) last and flatten them: (rows, n) view plus
    the outer shape (numbagg/utils.py:20-32)."""
    ndim = t.dim()
    moved = t.movedim(axes, tuple(range(ndim - len(axes), ndim))) if axes else t
    outer = tuple(moved.shape[: ndim - len(axes)])
    n = int(np.prod(moved.shape[ndim - len(axes):], dtype=np.int64))
    flat = moved.reshape(*outer, n) if outer else moved.reshape(n)
    v2, rows, n2, ld = dev.rows_view(flat)
    return v2, rows, n2, ld, outer


def _scalar_or_array(res: torch.Tensor, outer: tuple, as_tensor: bool):
    out = dev.to_output(res.view(outer) if outer else res.view(()), as_tensor)
    if not as_tensor and not outer:
        return out[()]  # NumPy gufuncs return a scalar for a 0-d result
    return out


class ndaggregate(_Op):
    """Plain aggregation over one or more axes (numbagg/decorators.py:188-272).

    ``axis=None`` reduces every axis; the reduced axes are moved last and
    flattened (numbagg/utils.py:20-32). Calls ``nbg_<name>(dtype, x, out, rows,
    n, ld, [ddof,] ws, ws_bytes, stream)``.
    """

    def __init__(self, func: Callable[..., Any], *, loops, out_dtype=None, supports_ddof: bool = False) -> None:
        super().__init__(func)
        self.loops = [np.dtype(d) for d in loops]
        self.out_dtype = None if out_dtype is None else np.dtype(out_dtype)
        self.supports_ddof = supports_ddof
        if not supports_ddof:
            params = [p for p in self.__signature__.parameters.values() if p.name != "ddof"]
            self.__signature__ = self.__signature__.replace(parameters=params)

    def __call__(self, *arrays, ddof: int = 1, axis: int | tuple[int, ...] | None = None):
        if not all(isinstance(a, (np.ndarray, torch.Tensor)) for a in arrays):
            raise TypeError(f"All positional arguments to {self} must be arrays: {arrays}")
        if len(arrays) != 1:
            raise TypeError(f"{self.__name__} takes one array, got {len(arrays)}")
        (arr,) = arrays
        ndim = arr.ndim
        if axis is None:
            axis = tuple(range(ndim))
        elif not isinstance(axis, tuple):
            axis = (axis,)
        axes = normalize_axis_tuple(axis, ndim)
        as_tensor = dev.is_tensor(arr)
        loop = _first_safe_loop(dev.np_dtype(arr), self.loops, self.__name__)
        t = dev.to_device(arr, loop)
        v2, rows, n, ld, outer = _reduce_rows(t, axes)
        odt = self.out_dtype or loop
        out = torch.empty(rows, dtype=dev.torch_dtype(odt), device=t.device)
        if rows:
            ws_bytes = _lib.query("nbg_reduce_workspace_bytes", rows, n, dev.nbg_code(v2))
            ws, ws_bytes = dev.workspace(ws_bytes, t.device)
            args = [ctypes.c_int(dev.nbg_code(v2)), dev.ptr(v2), dev.ptr(out), rows, n, ld]
            if self.supports_ddof:
                args.append(int(ddof))
            args += [dev.ptr(ws), ws_bytes, dev.stream_handle()]
            _lib.call(self.entry, *args)
        return _scalar_or_array(out, outer, as_tensor)


class ndreduce(_Op):
    """Reduction with an index or extreme result (numbagg/decorators.py:904-1028).

    nanmax/nanmin return the extreme in the input dtype (int64 for integers,
    as the reference's int64 signatures), nanargmax/nanargmin the int64 flat
    index over the reduced axes. Zero-size and all-NaN slices raise the
    reference's ValueError.
    """

    def __init__(self, func: Callable[..., Any], *, index: bool) -> None:
        super().__init__(func)
        self.index = index

    _LOOPS = [np.dtype(np.int32), np.dtype(np.int64), np.dtype(np.float32), np.dtype(np.float64)]

    def __call__(self, arr, *args, axis: int | tuple[int, ...] | None = None):
        if args:
            raise TypeError(f"{self.__name__} takes no extra positional arguments")
        if not isinstance(arr, (np.ndarray, torch.Tensor)):
            arr = np.asarray(arr)
        ndim = arr.ndim
        if axis is None:
            axes = tuple(range(ndim))
        elif isinstance(axis, (int, np.integer)):
            axes = normalize_axis_tuple(int(axis), ndim)
        else:
            axes = normalize_axis_tuple(tuple(axis), ndim)
        as_tensor = dev.is_tensor(arr)
        loop = _first_safe_loop(dev.np_dtype(arr), self._LOOPS, self.__name__)
        t = dev.to_device(arr, loop)
        v2, rows, n, ld, outer = _reduce_rows(t, axes)
        if self.index:
            odt = np.dtype(np.int64)
        else:
            odt = np.dtype(np.int64) if loop.kind == "i" else loop
        out = torch.empty(rows, dtype=dev.torch_dtype(odt), device=t.device)
        if rows and n == 0:
            if self.index:
                raise ValueError("All-NaN slice encountered")
            op = "fmax" if self.__name__ == "nanmax" else "fmin"
            raise ValueError(f"zero-size array to reduction operation {op} which has no identity")
        if rows:
            ws_bytes = _lib.query("nbg_reduce_workspace_bytes", rows, n, dev.nbg_code(v2))
            ws, ws_bytes = dev.workspace(ws_bytes, t.device)
            _lib.call(self.entry, ctypes.c_int(dev.nbg_code(v2)), dev.ptr(v2), dev.ptr(out), rows, n, ld,
                      dev.ptr(ws), ws_bytes, dev.stream_handle())
            if self.index and bool((out < 0).any().item()):
                raise ValueError("All-NaN slice encountered")
        return _scalar_or_array(out, outer, as_tensor)


def _matrix_loop(a, func_name: str) -> np.dtype:
    """Loop dtype of the float32/float64 matrix signatures for ``a``."""
    return _loop_dtype([dev.np_dtype(a)], func_name)


def _batched(t: torch.Tensor) -> tuple[torch.Tensor, tuple[int, ...], int]:
    """(batch, d0, d1) contiguous view of ``t`` (..., d0, d1)."""
    outer = tuple(t.shape[:-2])
    batch = int(np.prod(outer, dtype=np.int64)) if outer else 1
    return t.contiguous().view(batch, t.shape[-2], t.shape[-1]), outer, batch


class ndmatrix(_Op):
    """Static matrix operator (numbagg/decorators.py:677-744): fixed core
    dimensions (..., vars, obs) -> (..., vars, vars), no axis argument.

    Calls ``nbg_<name>(dtype, x, out, batch, nvars, nobs, ws, ws_bytes, stream)``.
    """

    def __call__(self, a, **kwargs):
        if "axis" in kwargs or "axes" in kwargs:
            raise TypeError(f"{self.__name__}() got an unexpected keyword argument 'axis'")
        dest = self._take_out(kwargs)
        if len(_shape(a)) < 2:
            raise ValueError(
                f"{self.__name__} requires at least a 2D array with shape (..., vars, obs). "
                "For 1D arrays, use nanvar for variance calculations."
            )
        as_tensor = dev.is_tensor(a)
        loop = _matrix_loop(a, self.__name__)
        t = dev.to_device(a, loop)
        x, outer, batch = _batched(t)
        nv, nobs = x.shape[1], x.shape[2]
        out = torch.empty((batch, nv, nv), dtype=dev.torch_dtype(loop), device=t.device)
        if batch and nv:
            if nobs == 0:  # no pair has an observation
                out.fill_(float("nan"))
            else:
                ws_bytes = _lib.query("nbg_matrix_workspace_bytes", batch, nv, nobs)
                ws, ws_bytes = dev.workspace(ws_bytes, t.device)
                _lib.call(self.entry, ctypes.c_int(dev.nbg_code(out)), dev.ptr(x), dev.ptr(out), batch, nv, nobs,
                          dev.ptr(ws), ws_bytes, dev.stream_handle())
        return self._deliver(dev.to_output(out.view(*outer, nv, nv), as_tensor), dest)


class ndmovematrix(_Op):
    """Moving-window matrix operator (numbagg/decorators.py:747-818): fixed core
    dimensions (..., obs, vars) -> (..., obs, vars, vars).

    Calls ``nbg_<name>(dtype, x, out, batch, nobs, nvars, window, min_count, ws,
    ws_bytes, stream)``.
    """

    def __call__(self, a, window: int, min_count: int | None = None, **kwargs):
        dest = self._take_out(kwargs)
        shape = _shape(a)
        if len(shape) < 2:
            raise ValueError(f"{self.__name__} requires at least a 2D array with shape (..., obs, vars).")
        if min_count is None:
            min_count = window
        elif min_count < 0:
            raise ValueError(f"min_count must be positive: {min_count}")
        if not 0 < window <= shape[-2]:
            raise ValueError(f"window not in valid range: {window}")
        if not _is_int_scalar(window) or not _is_int_scalar(min_count):
            raise TypeError(
                f"ufunc '{self.__name__}' not supported for the input types: window and min_count must be integers"
            )
        as_tensor = dev.is_tensor(a)
        loop = _matrix_loop(a, self.__name__)
        t = dev.to_device(a, loop)
        x, outer, batch = _batched(t)
        nobs, nv = x.shape[1], x.shape[2]
        out = torch.empty((batch, nobs, nv, nv), dtype=dev.torch_dtype(loop), device=t.device)
        if batch and nobs and nv:
            ws_bytes = _lib.query("nbg_move_matrix_workspace_bytes", batch, nobs, nv, int(window))
            ws, ws_bytes = dev.workspace(ws_bytes, t.device)
            _lib.call(self.entry, ctypes.c_int(dev.nbg_code(out)), dev.ptr(x), dev.ptr(out), batch, nobs, nv,
                      int(window), int(min_count), dev.ptr(ws), ws_bytes, dev.stream_handle())
        return self._deliver(dev.to_output(out.view(*outer, nobs, nv, nv), as_tensor), dest)


class ndmoveexpmatrix(_Op):
    """Exponentially weighted matrix operator (numbagg/decorators.py:1031-1112):
    fixed core dimensions (..., obs, vars), (obs) alpha -> (..., obs, vars, vars).

    A non-array alpha is broadcast along the observations (as a float64 array
    for a Python float, so float32 data runs on the float64 loop, as NumPy's
    loop selection does in the reference); an array alpha has its last axis
    along the observations and broadcasts over the leading dimensions.
    """

    def __call__(self, a, alpha, min_weight: float = 0, **kwargs):
        dest = self._take_out(kwargs)
        shape = _shape(a)
        if len(shape) < 2:
            raise ValueError(f"{self.__name__} requires at least a 2D array with shape (..., obs, vars).")
        as_tensor = dev.is_tensor(a)
        nobs = shape[-2]
        if isinstance(alpha, (np.ndarray, torch.Tensor)):
            alpha_dtype = dev.np_dtype(alpha)
            alpha_shape = tuple(alpha.shape)
        else:
            alpha_dtype = np.asarray(alpha).dtype
            alpha_shape = ()
        dtypes = [dev.np_dtype(a), alpha_dtype]
        if isinstance(min_weight, np.generic):
            dtypes.append(min_weight.dtype)
        loop = _loop_dtype(dtypes, self.__name__)
        mw = float(np.asarray(min_weight, dtype=loop))
        fused = loop == np.float64 and dev.np_dtype(a) == np.float32
        x_dtype = np.dtype(np.float32) if fused else loop
        t = dev.to_device(a, x_dtype)
        x, outer, batch = _batched(t)
        nv = x.shape[2]
        if alpha_shape and alpha_shape[-1] != nobs:
            raise ValueError(
                f"{self.__name__}: alpha has length {alpha_shape[-1]} but the observations dimension has "
                f"length {nobs}"
            )
        if len(alpha_shape) <= 1:
            a_row = torch.full((nobs,), float(np.asarray(alpha, dtype=loop)), dtype=dev.torch_dtype(loop),
                               device=t.device) if not alpha_shape else dev.to_device(alpha, loop).contiguous()
            alpha_t, ld_alpha = a_row, 0
        else:
            a_full = dev.to_device(alpha, loop)
            alpha_t = a_full.expand(*outer, nobs).contiguous().view(max(batch, 1), nobs)
            ld_alpha = nobs
        out = torch.empty((batch, nobs, nv, nv), dtype=dev.torch_dtype(loop), device=t.device)
        if batch and nobs and nv:
            code_in = _lib.NBG_F32 if x_dtype == np.float32 else _lib.NBG_F64
            ws_bytes = _lib.query("nbg_move_exp_matrix_workspace_bytes", batch, nobs, nv)
            ws, ws_bytes = dev.workspace(ws_bytes, t.device)
            _lib.call(self.entry, ctypes.c_int(code_in), ctypes.c_int(dev.nbg_code(out)), dev.ptr(x),
                      dev.ptr(alpha_t), ld_alpha, ctypes.c_double(mw), dev.ptr(out), batch, nobs, nv, dev.ptr(ws),
                      ws_bytes, dev.stream_handle())
        return self._deliver(dev.to_output(out.view(*outer, nobs, nv, nv), as_tensor), dest)
