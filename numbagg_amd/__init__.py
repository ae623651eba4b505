"""numbagg_amd: numbagg's hot path on AMD MI355X (gfx950) HIP kernels.

Same public names, signatures and argument semantics as numbagg for the grouped
reductions, the moving-window and exponentially weighted scans, and
ffill/bfill (numbagg/__init__.py). NumPy inputs are computed on the GPU and
returned as NumPy arrays; torch tensors on the GPU stay on the GPU. Also the
plain reductions, the correlation / covariance matrices and a device label
factorisation.
"""

from ._lib import spin_timeouts
from .factorize import factorize
from .funcs import (
    allnan,
    anynan,
    bfill,
    count,
    ffill,
    nanargmax,
    nanargmin,
    nancorrmatrix,
    nancount,
    nancovmatrix,
    nanmax,
    nanmean,
    nanmin,
    nanstd,
    nansum,
    nanvar,
)
from .grouped import (
    group_nanall,
    group_nanany,
    group_nanargmax,
    group_nanargmin,
    group_nancount,
    group_nanfirst,
    group_nanlast,
    group_nanmax,
    group_nanmean,
    group_nanmin,
    group_nanprod,
    group_nanstd,
    group_nansum,
    group_nansum_of_squares,
    group_nanvar,
)
from .moving import move_corr, move_cov, move_mean, move_std, move_sum, move_var
from .moving_matrix import move_corrmatrix, move_covmatrix, move_exp_nancorrmatrix, move_exp_nancovmatrix
from .moving_exp import (
    move_exp_nancorr,
    move_exp_nancount,
    move_exp_nancov,
    move_exp_nanmean,
    move_exp_nanstd,
    move_exp_nansum,
    move_exp_nanvar,
)

# Family lists restricted to the accelerated functions (numbagg/__init__.py:64-127).
GROUPED_FUNCS = [
    group_nanall,
    group_nanany,
    group_nanargmax,
    group_nanargmin,
    group_nancount,
    group_nanfirst,
    group_nanlast,
    group_nanmax,
    group_nanmean,
    group_nanmin,
    group_nanprod,
    group_nanstd,
    group_nansum,
    group_nansum_of_squares,
    group_nanvar,
]

MOVE_EXP_FUNCS = [
    move_exp_nancorr,
    move_exp_nancount,
    move_exp_nancov,
    move_exp_nanmean,
    move_exp_nanstd,
    move_exp_nansum,
    move_exp_nanvar,
]

MOVE_EXP_MATRIX_FUNCS = [move_exp_nancorrmatrix, move_exp_nancovmatrix]

MOVE_FUNCS = [move_corr, move_cov, move_mean, move_corrmatrix, move_covmatrix, move_std, move_sum, move_var]

AGGREGATION_FUNCS = [
    allnan,
    anynan,
    count,
    nanargmax,
    nanargmin,
    nancount,
    nanmax,
    nanmean,
    nanmin,
    nanstd,
    nansum,
    nanvar,
]

MATRIX_FUNCS = [nancorrmatrix, nancovmatrix]

OTHER_FUNCS = [bfill, ffill]

__version__ = "0.1.0"

__all__ = [
    "__version__",
    "allnan",
    "anynan",
    "bfill",
    "count",
    "factorize",
    "ffill",
    "nanargmax",
    "nanargmin",
    "nancorrmatrix",
    "nancount",
    "nancovmatrix",
    "nanmax",
    "nanmean",
    "nanmin",
    "nanstd",
    "nansum",
    "nanvar",
    "group_nanall",
    "group_nanany",
    "group_nanargmax",
    "group_nanargmin",
    "group_nancount",
    "group_nanfirst",
    "group_nanlast",
    "group_nanmax",
    "group_nanmean",
    "group_nanmin",
    "group_nanprod",
    "group_nanstd",
    "group_nansum",
    "group_nansum_of_squares",
    "group_nanvar",
    "move_corr",
    "move_corrmatrix",
    "move_cov",
    "move_covmatrix",
    "move_exp_nancorr",
    "move_exp_nancorrmatrix",
    "move_exp_nancount",
    "move_exp_nancov",
    "move_exp_nancovmatrix",
    "move_exp_nanmean",
    "move_exp_nanstd",
    "move_exp_nansum",
    "move_exp_nanvar",
    "move_mean",
    "move_std",
    "move_sum",
    "move_var",
    "spin_timeouts",
    "AGGREGATION_FUNCS",
    "GROUPED_FUNCS",
    "MATRIX_FUNCS",
    "MOVE_EXP_FUNCS",
    "MOVE_EXP_MATRIX_FUNCS",
    "MOVE_FUNCS",
    "OTHER_FUNCS",
]
