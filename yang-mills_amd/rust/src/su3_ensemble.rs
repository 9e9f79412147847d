//! Device counterparts of the reference's su3_ensemble.rs hot path:
//! `metropolis_sweep` (su3_ensemble.rs:342-391), `average_plaquette`
//! (:228-259), `run_chain` (:408-480) and `derive_su3_chain_seed` (:189-192),
//! with the reference's names, argument meaning and error variants.  The
//! sweep visits the links in even/odd checkerboard order with counter-based
//! random numbers, so the device chain is a different Markov chain from the
//! reference's lexicographic SplitMix64 chain with the same stationary
//! distribution.

use std::fmt;

use crate::ffi;
use crate::lattice::{diagnostics, ChainDiagnostics, DeviceLattice, DeviceStream, Group, Precision, Status,
                     YM_INVALID_ARGUMENT};

/// Proposals per link in one reference sweep: one per generator 0..7.
pub const SU3_GENERATORS: i32 = 8;
pub const SAMPLER_ID: &str = "YM-SU3-CHECKERBOARD-METROPOLIS-MI355X-V1";
pub const PROPOSAL_ID: &str = "YM-SU3-GENERATOR-METROPOLIS-V1";
pub const RNG_ID: &str = "PHILOX4X32-10";
pub const STREAM_ID: &str = "YM-SU3-COUNTER-SITE-MU-HIT-SWEEP-V1";

#[derive(Clone, Copy, Debug, Default, PartialEq, Eq)]
pub struct MetropolisAccounting {
    pub proposals: u64,
    pub accepts: u64,
}

/// The reference's `Su3EnsembleError` (su3_ensemble.rs:124-138) for the
/// statuses the device returns; anything else is `Device`.
#[derive(Clone, Copy, Debug, PartialEq, Eq)]
pub enum Su3EnsembleError {
    BetaOutsideReferenceRange,
    ProposalAngleOutsideReferenceRange,
    MeasurementCountOutsideReferenceRange,
    ZeroMeasurementStride,
    AutocorrelationLagOutOfRange,
    ScheduleOutsideReferenceRange,
    NonFiniteStatistic,
    Device(Status),
}

impl From<Status> for Su3EnsembleError {
    fn from(status: Status) -> Self {
        match status.0 {
            30 => Self::BetaOutsideReferenceRange,
            31 => Self::ProposalAngleOutsideReferenceRange,
            32 => Self::MeasurementCountOutsideReferenceRange,
            33 => Self::ZeroMeasurementStride,
            34 => Self::AutocorrelationLagOutOfRange,
            35 => Self::ScheduleOutsideReferenceRange,
            37 => Self::NonFiniteStatistic,
            _ => Self::Device(status),
        }
    }
}

impl fmt::Display for Su3EnsembleError {
    fn fmt(&self, f: &mut fmt::Formatter<'_>) -> fmt::Result {
        match self {
            Self::Device(status) => write!(f, "SU(3) device call failed: {}", status),
            other => write!(f, "SU(3) ensemble error: {:?}", other),
        }
    }
}

impl std::error::Error for Su3EnsembleError {}

pub fn derive_su3_chain_seed(master_seed: u64, chain_id: u64) -> u64 {
    unsafe { ffi::ym_derive_su3_chain_seed(master_seed, chain_id) }
}

/// su3_ensemble.rs:342-391 on device-resident links: one sweep, eight
/// single-generator proposals per link (generator h mod 8), angles on the
/// reference's paired grid, accept iff dS <= 0 or u < exp(-dS) in FP64.
/// `rng` advances by one sweep.
pub fn metropolis_sweep(lattice: &mut DeviceLattice, beta: f64, proposal_angle: f64, rng: &mut DeviceStream)
                        -> Result<MetropolisAccounting, Su3EnsembleError> {
    metropolis_sweeps(lattice, beta, proposal_angle, SU3_GENERATORS, 1, rng)
}

/// `n_sweeps` sweeps of `hits` proposals per link in one call (no host
/// synchronisation between the sweeps).
pub fn metropolis_sweeps(lattice: &mut DeviceLattice, beta: f64, proposal_angle: f64, hits: i32, n_sweeps: i32,
                         rng: &mut DeviceStream) -> Result<MetropolisAccounting, Su3EnsembleError> {
    if lattice.group() != Group::Su3 || n_sweeps < 0 {
        return Err(Su3EnsembleError::Device(Status(YM_INVALID_ARGUMENT)));
    }
    let mut accounting = MetropolisAccounting::default();
    Status::check(unsafe {
        ffi::ym_su3_metropolis_sweeps(lattice.raw(), beta, proposal_angle, hits, rng.key, rng.next_sweep,
                                      n_sweeps, 1, &mut accounting.proposals, &mut accounting.accepts)
    })?;
    rng.next_sweep += n_sweeps as u64;
    Ok(accounting)
}

/// su3_ensemble.rs:228-259: (1/6V) sum of (1/3) Re Tr of the plaquettes.
pub fn average_plaquette(lattice: &DeviceLattice) -> Result<f64, Su3EnsembleError> {
    let mut value = 0.0;
    Status::check(unsafe { ffi::ym_average_plaquette(lattice.raw(), &mut value) })?;
    Ok(value)
}

/// su3_ensemble.rs:36-46.
#[derive(Clone, Copy, Debug, PartialEq)]
pub struct Su3EnsembleConfig {
    pub extents: [usize; 4],
    pub beta: f64,
    pub master_seed: u64,
    pub chain_id: u64,
    pub thermalization_sweeps: usize,
    pub measurements: usize,
    pub sweeps_between_measurements: usize,
    pub proposal_angle: f64,
    pub autocorrelation_max_lag: usize,
}

/// Device options of a chain: storage precision, proposals per link per
/// sweep (8 = the reference) and the start (cold = the reference).
#[derive(Clone, Copy, Debug, PartialEq, Eq)]
pub struct DeviceOptions {
    pub precision: Precision,
    pub hits: i32,
    pub hot_start: bool,
}

impl Default for DeviceOptions {
    fn default() -> Self {
        Self { precision: Precision::Fp64, hits: SU3_GENERATORS, hot_start: false }
    }
}

/// su3_ensemble.rs:100-115.
#[derive(Clone, Debug, PartialEq)]
pub struct Su3ChainReport {
    pub sampler_id: &'static str,
    pub proposal_id: &'static str,
    pub rng_id: &'static str,
    pub stream_id: &'static str,
    pub config: Su3EnsembleConfig,
    pub derived_seed: u64,
    pub total_sweeps: usize,
    pub metropolis_proposals: u64,
    pub metropolis_accepts: u64,
    pub mean_metropolis_acceptance: f64,
    pub measurements: Vec<f64>,
    pub diagnostics: ChainDiagnostics,
    pub final_link_unitarity_residual: f64,
    pub final_link_determinant_residual: f64,
}

/// su3_ensemble.rs:408-480 on the device with FP64 storage and the
/// reference's eight proposals per link from a cold start.
pub fn run_chain(config: Su3EnsembleConfig) -> Result<Su3ChainReport, Su3EnsembleError> {
    run_chain_with(config, DeviceOptions::default())
}

pub fn run_chain_with(config: Su3EnsembleConfig, options: DeviceOptions)
                      -> Result<Su3ChainReport, Su3EnsembleError> {
    let c = ffi::YmSu3ChainConfig {
        extents: config.extents.map(|e| e as i64),
        beta: config.beta,
        master_seed: config.master_seed,
        chain_id: config.chain_id,
        thermalization_sweeps: config.thermalization_sweeps as i64,
        measurements: config.measurements as i64,
        sweeps_between_measurements: config.sweeps_between_measurements as i64,
        proposal_angle: config.proposal_angle,
        autocorrelation_max_lag: config.autocorrelation_max_lag as i64,
        hits: options.hits,
        precision: options.precision as i32,
        hot_start: options.hot_start as i32,
        reunitarize: 1,
    };
    let mut measurements = vec![0.0; config.measurements];
    let mut report = ffi::YmSu3ChainReport::default();
    Status::check(unsafe { ffi::ym_su3_run_chain(&c, measurements.as_mut_ptr(), &mut report) })?;
    let diagnostics = diagnostics(&measurements, config.autocorrelation_max_lag)?;
    Ok(Su3ChainReport {
        sampler_id: SAMPLER_ID,
        proposal_id: PROPOSAL_ID,
        rng_id: RNG_ID,
        stream_id: STREAM_ID,
        config,
        derived_seed: report.derived_seed,
        total_sweeps: report.total_sweeps as usize,
        metropolis_proposals: report.proposals,
        metropolis_accepts: report.accepts,
        mean_metropolis_acceptance: report.mean_acceptance,
        measurements,
        diagnostics,
        final_link_unitarity_residual: report.final_unitarity_residual,
        final_link_determinant_residual: report.final_determinant_residual,
    })
}
