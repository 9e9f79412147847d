"""ym_amd -- Python mirror of the reference lattice API over libym_amd.so.

The names follow the reference's Rust module functions (reference/rust/src/):
``Lattice4`` / ``Su3Lattice4`` handles with ``plaquette``-family observables,
``metropolis_sweep`` (su3_ensemble.rs:342), ``heatbath_sweep`` /
``overrelaxation_sweep`` (su2_ensemble.rs:321-346), ``average_plaquette``
(su2_ensemble.rs:348, su3_ensemble.rs:228), ``spatial_stout_smear``
(glueball_smearing.rs:257), ``spatial_stride2_block`` (glueball_blocking.rs:110), ``scalar_a1pp_operator_basis``
(glueball_operators.rs:143), ``run_chain`` (su3_ensemble.rs:408),
``su2_run_chain`` (su2_ensemble.rs:488) and
``diagnostics`` (su2_ensemble.rs:375).  Every call goes through the C ABI
declared in include/ym_amd.h; there is no CPU fallback: importing this package
on a machine without the built library raises, and device calls without a
GPU return YM_NO_DEVICE as a ``YmError``.
"""
from ._lib import YmError, lib, library_path, status_string  # noqa: F401
from . import stats  # noqa: F401
from . import ensemble  # noqa: F401
from .ensemble import Ensemble  # noqa: F401
from . import learned  # noqa: F401
from . import precision  # noqa: F401
from . import codec  # noqa: F401
from .lattice import (  # noqa: F401
    Lattice4,
    Su3Lattice4,
    ChainReport,
    Su2ChainReport,
    average_plaquette,
    blocked_extents,
    derive_chain_seed,
    derive_su3_chain_seed,
    diagnostics,
    heatbath_sweep,
    metropolis_sweep,
    overrelaxation_sweep,
    run_chain,
    scalar_a1pp_operator_basis,
    spatial_stout_smear,
    spatial_stride2_block,
    su2_run_chain,
)
