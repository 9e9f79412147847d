"""Self-checking chain runs on the device that print `key=value` evidence,
the device counterparts of the reference binaries
bin/su3_ensemble_reference.rs and bin/su2_ensemble_reference.rs: the same
configuration fields (defaults: the binaries' fixed configurations), the same
gates and the same report keys.  The device chains use their own sampler and
stream (Philox4x32-10 counters, checkerboard sub-passes), so the output names
them instead of the reference's SplitMix64 identifiers and claims no
reference contract.

    python -m ym_amd su3-ensemble [--extents 2,2,2,2 --beta 0 ...]
    python -m ym_amd su2-ensemble [--extents 2,2,2,2 --beta 0 ...]
"""
import argparse
import sys

import numpy as np

from . import lattice as _lat

SU3_SAMPLER_ID = "YM-SU3-DEVICE-CHECKERBOARD-V1"
SU3_PROPOSAL_ID = "YM-SU3-GENERATOR-METROPOLIS-V1 (generator h mod 8)"
RNG_ID = "PHILOX4X32-10"
SU3_STREAM_ID = "counter(site, mu | kind << 2 | index << 5, sweep)"
ANGLE_GRID_MAGNITUDE_BITS = 52          # su3_ensemble.rs:31
MAX_REFERENCE_SWEEPS = 512              # su3_ensemble.rs:25
LINK_RESIDUAL_LIMIT = 5.0e-12           # su3_ensemble.rs:30 (FP64 links)
FP32_LINK_RESIDUAL_LIMIT = 1.0e-6       # FP32 storage after FP64 reunitarisation


class GateFailure(RuntimeError):
    pass


def _series(label, values):
    return f"{label}=" + ",".join(f"{v:.17e}" for v in values)


def _extents(text):
    ext = tuple(int(v) for v in text.split(","))
    if len(ext) != 4:
        raise argparse.ArgumentTypeError("four comma-separated extents")
    return ext


def su3_ensemble(extents=(2, 2, 2, 2), beta=0.0, master_seed=0x3C00198212230101, chain_id=0,
                 thermalization_sweeps=4, measurements=16, sweeps_between_measurements=1, proposal_angle=0.35,
                 autocorrelation_max_lag=4, hits=8, precision=64):
    """run_chain (su3_ensemble.rs:408-480) on the device plus the gates of
    bin/su3_ensemble_reference.rs:45-59; returns the `key=value` lines."""
    r = _lat.run_chain(extents, beta, master_seed, chain_id, thermalization_sweeps, measurements,
                       sweeps_between_measurements, proposal_angle, autocorrelation_max_lag, hits=hits,
                       precision=precision)
    limit = LINK_RESIDUAL_LIMIT if precision == 64 else FP32_LINK_RESIDUAL_LIMIT
    if beta == 0.0 and r.mean_acceptance != 1.0:
        raise GateFailure("beta=0 SU(3) Metropolis sweep must accept every proposal")
    if r.final_unitarity_residual > limit or r.final_determinant_residual > limit:
        raise GateFailure("SU(3) link residual exceeded the limit")
    if np.any((r.measurements < -0.5 - 1e-12) | (r.measurements > 1.0 + 1e-12)):
        raise GateFailure("SU(3) normalized plaquette left its physical trace range")
    d, ac = _lat.diagnostics(r.measurements, autocorrelation_max_lag)
    return [
        "device_contract=YM-SU3-ENSEMBLE-DEVICE-V1",
        f"sampler_id={SU3_SAMPLER_ID}",
        f"proposal_id={SU3_PROPOSAL_ID}",
        f"rng_id={RNG_ID}",
        f"stream_id={SU3_STREAM_ID}",
        f"master_seed={master_seed:016x}",
        f"chain_id={chain_id}",
        f"derived_seed={r.derived_seed:016x}",
        "lattice=" + ",".join(str(e) for e in extents),
        f"link_precision=fp{precision}",
        f"beta={beta:.17e}",
        f"proposal_angle={proposal_angle:.17e}",
        f"hits_per_link={hits}",
        f"angle_grid_magnitude_bits={ANGLE_GRID_MAGNITUDE_BITS}",
        f"maximum_reference_sweeps={MAX_REFERENCE_SWEEPS}",
        f"thermalization_sweeps={thermalization_sweeps}",
        f"measurements={measurements}",
        f"sweeps_between_measurements={sweeps_between_measurements}",
        f"autocorrelation_max_lag={autocorrelation_max_lag}",
        f"total_sweeps={r.total_sweeps}",
        f"metropolis_proposals={r.proposals}",
        f"metropolis_accepts={r.accepts}",
        f"metropolis_acceptance={r.mean_acceptance:.17e}",
        _series("measurements", r.measurements),
        f"mean_plaquette={d.mean:.17e}",
        f"variance={d.variance:.17e}",
        _series("autocorrelation", ac),
        f"positive_window_lag={d.positive_window_lag}",
        f"tau_int={d.tau_int:.17e}",
        f"effective_sample_size={d.ess:.17e}",
        f"standard_error={d.standard_error:.17e}",
        f"final_link_unitarity_residual={r.final_unitarity_residual:.17e}",
        f"final_link_determinant_residual={r.final_determinant_residual:.17e}",
        f"link_residual_limit={limit:.17e}",
        "published_su3_baseline_reproduced=false",
        "glueball_analysis_enabled=false",
        "confinement_claimed=false",
        "continuum_qft_constructed=false",
        "positive_mass_gap_proved=false",
        "status=passed",
    ]


def _su2_chain(extents, beta, master_seed, chain_id, thermalization_sweeps, measurements,
               sweeps_between_measurements, overrelaxation_sweeps_per_heatbath, max_attempts, precision):
    """su2_ensemble.rs run_chain on the device (ym_su2_run_chain): cold start,
    Markov sweeps of one heat-bath sweep then `overrelaxation_sweeps_per_heatbath`
    over-relaxation sweeps (su2_ensemble.rs:465-486), the plaquette after
    every `sweeps_between_measurements` sweeps."""
    rep = _lat.su2_run_chain(extents, beta, master_seed, chain_id, thermalization_sweeps, measurements,
                             sweeps_between_measurements, overrelaxation_sweeps_per_heatbath, max_attempts,
                             autocorrelation_max_lag=1, precision=precision)
    return rep.derived_seed, rep.measurements, rep.mean_heatbath_acceptance, rep.final_link_unit_residual


def su2_ensemble(extents=(2, 2, 2, 2), beta=0.0, master_seed=0x0123456789ABCDEF, chain_id=0,
                 thermalization_sweeps=1, measurements=128, sweeps_between_measurements=1,
                 overrelaxation_sweeps_per_heatbath=1, max_attempts=10_000, autocorrelation_max_lag=16,
                 precision=64):
    """The SU(2) chain of bin/su2_ensemble_reference.rs on the device with its
    gates (deterministic replay, beta=0 Haar band, ESS, heat-bath acceptance,
    unit-link invariant, an independent chain id, over-relaxation preserving
    the action); returns the `key=value` lines."""
    args = (extents, beta, master_seed, chain_id, thermalization_sweeps, measurements,
            sweeps_between_measurements, overrelaxation_sweeps_per_heatbath, max_attempts, precision)
    key, meas, acc, unit = _su2_chain(*args)
    key2, meas2, _, _ = _su2_chain(*args)
    if key2 != key or not np.array_equal(meas, meas2):
        raise GateFailure("fixed chain did not replay deterministically")
    d, _ = _lat.diagnostics(meas, autocorrelation_max_lag)
    if beta == 0.0 and abs(d.mean) > 0.08:
        raise GateFailure("beta=0 average plaquette is outside the analytic Haar sanity band")
    if d.ess < 16.0:
        raise GateFailure("effective sample size is unexpectedly small")
    if not 0.5 <= acc <= 1.0:
        raise GateFailure("heat-bath scalar rejection acceptance is outside the sanity band")
    limit = 1.0e-12 if precision == 64 else FP32_LINK_RESIDUAL_LIMIT
    if unit > limit:
        raise GateFailure("ensemble generation violated the SU(2) unit-link invariant")
    key_i, meas_i, _, _ = _su2_chain(extents, beta, master_seed, chain_id + 1, *args[4:])
    if key_i == key or np.array_equal(meas_i, meas):
        raise GateFailure("independent chain identity did not produce an independent stream")
    # over-relaxation is microcanonical: action before and after one sweep at beta = 2
    lat = _lat.Lattice4((2, 2, 2, 2), 64)
    lat.heatbath_sweeps(2.0, _lat.derive_chain_seed(0xFEEDFACECAFEBEEF, 0), 0, 1, 100_000)
    before = lat.wilson_action(2.0)
    lat.overrelax_sweeps(1)
    residual = abs(lat.wilson_action(2.0) - before)
    if residual > 2.0e-10:
        raise GateFailure("over-relaxation failed the microcanonical action check")
    return [
        "device_contract=YM-SU2-ENSEMBLE-DEVICE-V1",
        f"rng={RNG_ID}",
        f"derived_seed={key:016x}",
        "lattice=" + ",".join(str(e) for e in extents),
        f"link_precision=fp{precision}",
        f"beta0_mean_plaquette={d.mean:.17e}" if beta == 0.0 else f"mean_plaquette={d.mean:.17e}",
        f"beta0_tau_int={d.tau_int:.17e}" if beta == 0.0 else f"tau_int={d.tau_int:.17e}",
        (f"beta0_effective_sample_size={d.ess:.17e}" if beta == 0.0 else f"effective_sample_size={d.ess:.17e}"),
        f"heatbath_acceptance={acc:.17e}",
        f"overrelaxation_action_residual={residual:.17e}",
        f"final_link_unit_residual={unit:.17e}",
        "published_plaquette_baseline_reproduced=false",
        "confinement_claimed=false",
        "continuum_mass_gap_claimed=false",
        "status=passed",
    ]


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m ym_amd", description=__doc__.split("\n\n")[0])
    sub = ap.add_subparsers(dest="cmd", required=True)
    s3 = sub.add_parser("su3-ensemble", help="device run_chain with the gates of su3_ensemble_reference")
    s3.add_argument("--extents", type=_extents, default=(2, 2, 2, 2))
    s3.add_argument("--beta", type=float, default=0.0)
    s3.add_argument("--master-seed", type=lambda v: int(v, 0), default=0x3C00198212230101)
    s3.add_argument("--chain-id", type=int, default=0)
    s3.add_argument("--thermalization-sweeps", type=int, default=4)
    s3.add_argument("--measurements", type=int, default=16)
    s3.add_argument("--sweeps-between-measurements", type=int, default=1)
    s3.add_argument("--proposal-angle", type=float, default=0.35)
    s3.add_argument("--autocorrelation-max-lag", type=int, default=4)
    s3.add_argument("--hits", type=int, default=8)
    s3.add_argument("--precision", type=int, choices=[32, 64], default=64)
    s2 = sub.add_parser("su2-ensemble", help="device SU(2) chain with the gates of su2_ensemble_reference")
    s2.add_argument("--extents", type=_extents, default=(2, 2, 2, 2))
    s2.add_argument("--beta", type=float, default=0.0)
    s2.add_argument("--master-seed", type=lambda v: int(v, 0), default=0x0123456789ABCDEF)
    s2.add_argument("--chain-id", type=int, default=0)
    s2.add_argument("--thermalization-sweeps", type=int, default=1)
    s2.add_argument("--measurements", type=int, default=128)
    s2.add_argument("--sweeps-between-measurements", type=int, default=1)
    s2.add_argument("--overrelaxation-sweeps-per-heatbath", type=int, default=1)
    s2.add_argument("--max-attempts", type=int, default=10_000)
    s2.add_argument("--autocorrelation-max-lag", type=int, default=16)
    s2.add_argument("--precision", type=int, choices=[32, 64], default=64)
    a = ap.parse_args(argv)
    kw = {k: v for k, v in vars(a).items() if k != "cmd"}
    try:
        lines = su3_ensemble(**kw) if a.cmd == "su3-ensemble" else su2_ensemble(**kw)
    except GateFailure as e:
        print(f"status=failed\nreason={e}", file=sys.stderr)
        return 1
    print("\n".join(lines))
    return 0
