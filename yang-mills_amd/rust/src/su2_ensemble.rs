//! Device counterparts of the reference's su2_ensemble.rs hot path:
//! `heatbath_sweep` (su2_ensemble.rs:321-342), `overrelaxation_sweep`
//! (:344-346), `markov_sweep` (:465-486), `average_plaquette` (:348-373),
//! `run_chain` (:488-552) and `derive_chain_seed` (:52-55), with the
//! reference's names, argument meaning and error variants.  Sweeps visit the
//! links in even/odd checkerboard order; attempt a of a link draws
//! Philox(site, mu, a, sweep), so a link ends with the attempt, r0 and
//! proposal count of its own sequential rejection loop.

use std::fmt;

use crate::ffi;
use crate::lattice::{diagnostics, ChainDiagnostics, DeviceLattice, DeviceStream, Group, Precision, Status,
                     YM_INVALID_ARGUMENT};

pub const SAMPLER_ID: &str = "YM-SU2-CHECKERBOARD-ENSEMBLE-MI355X-V1";
pub const RNG_ID: &str = "PHILOX4X32-10";

/// su2_ensemble.rs:56-68.
#[derive(Clone, Copy, Debug, PartialEq)]
pub struct EnsembleConfig {
    pub extents: [usize; 4],
    pub beta: f64,
    pub master_seed: u64,
    pub chain_id: u64,
    pub thermalization_sweeps: usize,
    pub measurements: usize,
    pub sweeps_between_measurements: usize,
    pub overrelaxation_sweeps_per_heatbath: usize,
    pub max_heatbath_attempts_per_link: usize,
    pub autocorrelation_max_lag: usize,
}

/// The reference's `EnsembleError` (su2_ensemble.rs:110-121) for the
/// statuses the device returns; anything else is `Device`.
#[derive(Clone, Copy, Debug, PartialEq, Eq)]
pub enum EnsembleError {
    BetaOutsideReferenceRange,
    ScheduleOutsideReferenceRange,
    TooFewMeasurements,
    ZeroMeasurementStride,
    TooManyOverrelaxationSweeps,
    ZeroHeatbathAttemptLimit,
    AutocorrelationLagOutOfRange,
    HeatbathAttemptLimitExceeded,
    NonFiniteStatistic,
    Device(Status),
}

impl From<Status> for EnsembleError {
    fn from(status: Status) -> Self {
        match status.0 {
            30 => Self::BetaOutsideReferenceRange,
            35 => Self::ScheduleOutsideReferenceRange,
            32 => Self::TooFewMeasurements,
            33 => Self::ZeroMeasurementStride,
            40 => Self::TooManyOverrelaxationSweeps,
            41 => Self::ZeroHeatbathAttemptLimit,
            34 => Self::AutocorrelationLagOutOfRange,
            42 => Self::HeatbathAttemptLimitExceeded,
            37 => Self::NonFiniteStatistic,
            _ => Self::Device(status),
        }
    }
}

impl fmt::Display for EnsembleError {
    fn fmt(&self, f: &mut fmt::Formatter<'_>) -> fmt::Result {
        match self {
            Self::Device(status) => write!(f, "SU(2) device call failed: {}", status),
            other => write!(f, "SU(2) ensemble error: {:?}", other),
        }
    }
}

impl std::error::Error for EnsembleError {}

pub fn derive_chain_seed(master_seed: u64, chain_id: u64) -> u64 {
    unsafe { ffi::ym_derive_chain_seed(master_seed, chain_id) }
}

fn attempts_arg(max_attempts_per_link: usize) -> i32 {
    // the device takes an int32 limit: larger limits are clamped to i32::MAX
    max_attempts_per_link.min(i32::MAX as usize) as i32
}

fn require_su2(lattice: &DeviceLattice) -> Result<(), EnsembleError> {
    if lattice.group() == Group::Su2 {
        Ok(())
    } else {
        Err(EnsembleError::Device(Status(YM_INVALID_ARGUMENT)))
    }
}

/// su2_ensemble.rs:321-342: one heat-bath sweep; returns (attempts, updated
/// links).  A link that exhausts its attempts keeps its value and the sweep
/// returns `HeatbathAttemptLimitExceeded` after finishing.
pub fn heatbath_sweep(lattice: &mut DeviceLattice, beta: f64, rng: &mut DeviceStream, max_attempts_per_link: usize)
                      -> Result<(u64, u64), EnsembleError> {
    require_su2(lattice)?;
    let (mut proposals, mut updates) = (0u64, 0u64);
    let status = unsafe {
        ffi::ym_su2_heatbath_sweeps(lattice.raw(), beta, rng.key, rng.next_sweep, 1,
                                    attempts_arg(max_attempts_per_link), &mut proposals, &mut updates)
    };
    if status == 0 || status == 42 {
        rng.next_sweep += 1;
    }
    Status::check(status)?;
    Ok((proposals, updates))
}

/// su2_ensemble.rs:344-346: one over-relaxation sweep (no random numbers).
pub fn overrelaxation_sweep(lattice: &mut DeviceLattice) -> Result<(), EnsembleError> {
    require_su2(lattice)?;
    Status::check(unsafe { ffi::ym_su2_overrelax_sweeps(lattice.raw(), 1) })?;
    Ok(())
}

/// su2_ensemble.rs:465-486: `n_sweeps` x (one heat-bath sweep, then
/// `overrelaxation_sweeps_per_heatbath` over-relaxation sweeps) in one call;
/// returns (attempts, updated links).
pub fn markov_sweeps(lattice: &mut DeviceLattice, beta: f64, rng: &mut DeviceStream, n_sweeps: usize,
                     overrelaxation_sweeps_per_heatbath: usize, max_attempts_per_link: usize)
                     -> Result<(u64, u64), EnsembleError> {
    require_su2(lattice)?;
    let n = i32::try_from(n_sweeps).map_err(|_| EnsembleError::ScheduleOutsideReferenceRange)?;
    let or = i32::try_from(overrelaxation_sweeps_per_heatbath).map_err(|_| EnsembleError::TooManyOverrelaxationSweeps)?;
    let (mut proposals, mut updates) = (0u64, 0u64);
    let status = unsafe {
        ffi::ym_su2_markov_sweeps(lattice.raw(), beta, rng.key, rng.next_sweep, n, or,
                                  attempts_arg(max_attempts_per_link), &mut proposals, &mut updates)
    };
    if status == 0 || status == 42 {
        rng.next_sweep += n_sweeps as u64;
    }
    Status::check(status)?;
    Ok((proposals, updates))
}

/// su2_ensemble.rs:348-373: (1/6V) sum of (1/2) Re Tr of the plaquettes.
pub fn average_plaquette(lattice: &DeviceLattice) -> Result<f64, EnsembleError> {
    let mut value = 0.0;
    Status::check(unsafe { ffi::ym_average_plaquette(lattice.raw(), &mut value) })?;
    Ok(value)
}

/// Device options of a chain: storage precision and start (cold = the
/// reference).
#[derive(Clone, Copy, Debug, PartialEq, Eq)]
pub struct DeviceOptions {
    pub precision: Precision,
    pub hot_start: bool,
}

impl Default for DeviceOptions {
    fn default() -> Self {
        Self { precision: Precision::Fp64, hot_start: false }
    }
}

/// su2_ensemble.rs:173-186.
#[derive(Clone, Debug, PartialEq)]
pub struct ChainReport {
    pub sampler_id: &'static str,
    pub rng_id: &'static str,
    pub config: EnsembleConfig,
    pub derived_seed: u64,
    pub total_heatbath_sweeps: usize,
    pub total_overrelaxation_sweeps: usize,
    pub heatbath_proposals: u64,
    pub heatbath_updates: u64,
    pub mean_heatbath_acceptance: f64,
    pub measurements: Vec<f64>,
    pub diagnostics: ChainDiagnostics,
    pub final_link_unit_residual: f64,
}

/// su2_ensemble.rs:488-552 on the device (EnsembleConfig::validate's bounds
/// and order, su2_ensemble.rs:70-103), FP64 storage, cold start.
pub fn run_chain(config: EnsembleConfig) -> Result<ChainReport, EnsembleError> {
    run_chain_with(config, DeviceOptions::default())
}

pub fn run_chain_with(config: EnsembleConfig, options: DeviceOptions) -> Result<ChainReport, EnsembleError> {
    let c = ffi::YmSu2ChainConfig {
        extents: config.extents.map(|e| e as i64),
        beta: config.beta,
        master_seed: config.master_seed,
        chain_id: config.chain_id,
        thermalization_sweeps: config.thermalization_sweeps as i64,
        measurements: config.measurements as i64,
        sweeps_between_measurements: config.sweeps_between_measurements as i64,
        overrelaxation_sweeps_per_heatbath: config.overrelaxation_sweeps_per_heatbath as i64,
        max_heatbath_attempts_per_link: config.max_heatbath_attempts_per_link as i64,
        autocorrelation_max_lag: config.autocorrelation_max_lag as i64,
        precision: options.precision as i32,
        hot_start: options.hot_start as i32,
    };
    let mut measurements = vec![0.0; config.measurements];
    let mut report = ffi::YmSu2ChainReport::default();
    Status::check(unsafe { ffi::ym_su2_run_chain(&c, measurements.as_mut_ptr(), &mut report) })?;
    let diagnostics = diagnostics(&measurements, config.autocorrelation_max_lag)?;
    Ok(ChainReport {
        sampler_id: SAMPLER_ID,
        rng_id: RNG_ID,
        config,
        derived_seed: report.derived_seed,
        total_heatbath_sweeps: report.total_heatbath_sweeps as usize,
        total_overrelaxation_sweeps: report.total_overrelaxation_sweeps as usize,
        heatbath_proposals: report.heatbath_proposals,
        heatbath_updates: report.heatbath_updates,
        mean_heatbath_acceptance: report.mean_heatbath_acceptance,
        measurements,
        diagnostics,
        final_link_unit_residual: report.final_link_unit_residual,
    })
}
