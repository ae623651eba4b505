"""ctypes binding of ``libnumbagg_amd.so`` (the C ABI in include/numbagg_amd.h).

The product path has no CPU fallback: if the HIP library is missing or no ROCm
GPU is visible, every call raises ``RuntimeError``.
"""

from __future__ import annotations

import ctypes
import pathlib
import re
import threading

from ._build import LIB_PATH, ROOT_DIR

HEADER = ROOT_DIR / "include" / "numbagg_amd.h"

NBG_F32, NBG_F64, NBG_I32, NBG_I64 = 0, 1, 2, 3

_c_i64 = ctypes.c_int64
_c_int = ctypes.c_int
_c_ptr = ctypes.c_void_p
_c_size = ctypes.c_size_t
_c_dbl = ctypes.c_double

# Signatures of every entry point, mirroring include/numbagg_amd.h.
_SIGNATURES: dict[str, tuple[object, list[object]]] = {
    "nbg_version": (ctypes.c_char_p, []),
    "nbg_last_error": (ctypes.c_char_p, []),
    "nbg_spin_timeouts": (_c_int, [ctypes.POINTER(ctypes.c_uint)]),
    "nbg_reset_spin_timeouts": (_c_int, []),
    "nbg_stream_copy": (_c_int, [_c_ptr, _c_ptr, _c_size, _c_ptr]),
    "nbg_tune_set": (_c_int, [ctypes.c_char_p, _c_i64]),
    "nbg_tune_clear": (_c_int, []),
    "nbg_tune_get": (_c_i64, [ctypes.c_char_p, _c_i64]),
    "nbg_fill_workspace_bytes": (_c_size, [_c_i64, _c_i64, _c_int]),
    "nbg_ffill": (
        _c_int,
        [_c_int, _c_ptr, _c_ptr, _c_i64, _c_i64, _c_i64, _c_i64, _c_i64, _c_ptr, _c_size, _c_ptr],
    ),
    "nbg_bfill": (
        _c_int,
        [_c_int, _c_ptr, _c_ptr, _c_i64, _c_i64, _c_i64, _c_i64, _c_i64, _c_ptr, _c_size, _c_ptr],
    ),
}

_MOVE1 = (_c_int, [_c_int, _c_ptr, _c_ptr, _c_i64, _c_i64, _c_i64, _c_i64, _c_i64, _c_i64, _c_ptr])
_MOVE2 = (
    _c_int,
    [_c_int, _c_ptr, _c_ptr, _c_ptr, _c_i64, _c_i64, _c_i64, _c_i64, _c_i64, _c_i64, _c_i64, _c_ptr],
)
for _name in ("nbg_move_sum", "nbg_move_mean", "nbg_move_std", "nbg_move_var"):
    _SIGNATURES[_name] = _MOVE1
for _name in ("nbg_move_cov", "nbg_move_corr"):
    _SIGNATURES[_name] = _MOVE2
_SIGNATURES["nbg_move_exp_workspace_bytes"] = (_c_size, [_c_i64, _c_i64, _c_int])
_EXP1 = (
    _c_int,
    [_c_int, _c_int, _c_ptr, _c_ptr, _c_i64, _c_dbl, _c_ptr, _c_i64, _c_i64, _c_i64, _c_i64, _c_dbl,
     _c_ptr, _c_size, _c_ptr],
)
_EXP2 = (
    _c_int,
    [_c_int, _c_int, _c_ptr, _c_ptr, _c_ptr, _c_i64, _c_dbl, _c_ptr, _c_i64, _c_i64, _c_i64, _c_i64,
     _c_i64, _c_dbl, _c_ptr, _c_size, _c_ptr],
)
for _name in ("nancount", "nanmean", "nansum", "nanvar", "nanstd"):
    _SIGNATURES[f"nbg_move_exp_{_name}"] = _EXP1
for _name in ("nancov", "nancorr"):
    _SIGNATURES[f"nbg_move_exp_{_name}"] = _EXP2
_SIGNATURES["nbg_group_workspace_bytes"] = (_c_size, [_c_int, _c_i64, _c_i64, _c_i64, _c_int])
_GROUP = (
    _c_int,
    [_c_int, _c_ptr, _c_int, _c_ptr, _c_ptr, _c_i64, _c_i64, _c_i64, _c_i64, _c_int, _c_ptr, _c_size, _c_ptr],
)
_GROUP_DDOF = (
    _c_int,
    [_c_int, _c_ptr, _c_int, _c_ptr, _c_ptr, _c_i64, _c_i64, _c_i64, _c_i64, _c_i64, _c_int, _c_ptr, _c_size,
     _c_ptr],
)
for _name in ("nansum", "nanmean", "nancount", "nanargmax", "nanargmin", "nanfirst", "nanlast", "nanmin",
              "nanmax", "nanprod", "nansum_of_squares", "nanany", "nanall"):
    _SIGNATURES[f"nbg_group_{_name}"] = _GROUP
for _name in ("nanvar", "nanstd"):
    _SIGNATURES[f"nbg_group_{_name}"] = _GROUP_DDOF
_SIGNATURES["nbg_group_nansum_noalloc"] = (
    _c_int,
    [_c_int, _c_ptr, _c_int, _c_ptr, _c_ptr, _c_i64, _c_i64, _c_i64, _c_i64, _c_int, _c_ptr],
)
_SIGNATURES["nbg_reduce_workspace_bytes"] = (_c_size, [_c_i64, _c_i64, _c_int])
_RED = (_c_int, [_c_int, _c_ptr, _c_ptr, _c_i64, _c_i64, _c_i64, _c_ptr, _c_size, _c_ptr])
_RED_DDOF = (_c_int, [_c_int, _c_ptr, _c_ptr, _c_i64, _c_i64, _c_i64, _c_i64, _c_ptr, _c_size, _c_ptr])
for _name in ("nansum", "nanmean", "nancount", "allnan", "anynan", "nanmax", "nanmin", "nanargmax", "nanargmin"):
    _SIGNATURES[f"nbg_{_name}"] = _RED
for _name in ("nanvar", "nanstd"):
    _SIGNATURES[f"nbg_{_name}"] = _RED_DDOF
_SIGNATURES["nbg_factorize_workspace_bytes"] = (_c_size, [_c_i64])
_SIGNATURES["nbg_factorize_count"] = (_c_int, [_c_int, _c_ptr, _c_i64, _c_int, _c_ptr, _c_size, _c_ptr, _c_ptr])
_SIGNATURES["nbg_factorize_finish"] = (
    _c_int,
    [_c_int, _c_ptr, _c_i64, _c_i64, _c_int, _c_ptr, _c_ptr, _c_ptr, _c_size, _c_ptr],
)
_SIGNATURES["nbg_matrix_workspace_bytes"] = (_c_size, [_c_i64, _c_i64, _c_i64])
for _name in ("nancorrmatrix", "nancovmatrix"):
    _SIGNATURES[f"nbg_{_name}"] = (_c_int, [_c_int, _c_ptr, _c_ptr, _c_i64, _c_i64, _c_i64, _c_ptr, _c_size, _c_ptr])
_SIGNATURES["nbg_move_matrix_workspace_bytes"] = (_c_size, [_c_i64, _c_i64, _c_i64, _c_i64])
for _name in ("move_corrmatrix", "move_covmatrix"):
    _SIGNATURES[f"nbg_{_name}"] = (
        _c_int,
        [_c_int, _c_ptr, _c_ptr, _c_i64, _c_i64, _c_i64, _c_i64, _c_i64, _c_ptr, _c_size, _c_ptr],
    )
_SIGNATURES["nbg_move_exp_matrix_workspace_bytes"] = (_c_size, [_c_i64, _c_i64, _c_i64])
for _name in ("move_exp_nancorrmatrix", "move_exp_nancovmatrix"):
    _SIGNATURES[f"nbg_{_name}"] = (
        _c_int,
        [_c_int, _c_int, _c_ptr, _c_ptr, _c_i64, ctypes.c_double, _c_ptr, _c_i64, _c_i64, _c_i64, _c_ptr, _c_size,
         _c_ptr],
    )
_SIGNATURES["nbg_ragged_tile_size"] = (_c_i64, [])
_SIGNATURES["nbg_ragged_workspace_bytes"] = (_c_size, [_c_i64, _c_i64])
_SIGNATURES["nbg_ragged_max_lds_labels"] = (_c_i64, [])
_SIGNATURES["nbg_ragged_move_mean_group_nanmean"] = (
    _c_int,
    [_c_ptr, _c_ptr, _c_ptr, _c_i64, _c_ptr, _c_i64, _c_ptr, _c_ptr, _c_i64, _c_ptr, _c_i64, _c_i64, _c_i64,
     _c_ptr, _c_ptr, _c_ptr, _c_size, _c_ptr],
)
_c_flt = ctypes.c_float
_SIGNATURES["nbg_dense_forward"] = (_c_int, [_c_ptr, _c_ptr, _c_ptr, _c_ptr, _c_i64, _c_i64, _c_i64, _c_int, _c_ptr])
_SIGNATURES["nbg_dense_backward"] = (
    _c_int,
    [_c_ptr, _c_ptr, _c_ptr, _c_ptr, _c_ptr, _c_ptr, _c_ptr, _c_ptr, _c_i64, _c_i64, _c_i64, _c_int, _c_ptr],
)
_SIGNATURES["nbg_mse_grad"] = (_c_int, [_c_ptr, _c_ptr, _c_ptr, _c_ptr, _c_i64, _c_ptr])
_SIGNATURES["nbg_adam_step"] = (
    _c_int,
    [_c_ptr, _c_ptr, _c_ptr, _c_ptr, _c_i64, _c_flt, _c_flt, _c_flt, _c_flt, _c_i64, _c_ptr],
)

_lock = threading.Lock()
_lib: ctypes.CDLL | None = None


def header_symbols(header: pathlib.Path = HEADER) -> list[str]:
    """Names of the functions declared in the public header."""
    text = header.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(nbg_[a-z0-9_]+)\s*\(", text)))


def load() -> ctypes.CDLL:
    """Load (once) and type the HIP library; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            if not LIB_PATH.exists():
                raise RuntimeError(
                    f"numbagg_amd: HIP library {LIB_PATH} is not built; "
                    "run `python -c 'import __graft_entry__ as g; g.build()'`"
                )
            lib = ctypes.CDLL(str(LIB_PATH))
            for name, (restype, argtypes) in _SIGNATURES.items():
                fn = getattr(lib, name)
                fn.restype = restype
                fn.argtypes = argtypes
            _lib = lib
    return _lib


def last_error() -> str:
    msg = load().nbg_last_error()
    return msg.decode() if msg else ""


def call(name: str, *args) -> None:
    """Call an NBG entry point and raise ``RuntimeError`` on a non-zero status."""
    fn = getattr(load(), name)
    rc = fn(*args)
    if rc != 0:
        raise RuntimeError(f"{name} failed (status {rc}): {last_error()}")


def query(name: str, *args) -> int:
    return int(getattr(load(), name)(*args))


def spin_timeouts() -> int:
    c = ctypes.c_uint(0)
    call("nbg_spin_timeouts", ctypes.byref(c))
    return int(c.value)
