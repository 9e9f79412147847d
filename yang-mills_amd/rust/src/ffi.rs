//! Raw declarations of the include/ym_amd.h entry points this crate wraps.
//! Every signature is the header's, argument for argument; the header's
//! comments give the reference function each one replaces.

use std::os::raw::c_char;

/// `ym_lattice`: one device link field, its stream and its scratch buffers.
#[repr(C)]
pub struct YmLattice {
    _opaque: [u8; 0],
}

/// `ym_chain_diagnostics` (su2_ensemble.rs:162-170 without the vector).
#[repr(C)]
#[derive(Clone, Copy, Debug, Default, PartialEq)]
pub struct YmChainDiagnostics {
    pub mean: f64,
    pub variance: f64,
    pub positive_window_lag: i64,
    pub tau_int: f64,
    pub ess: f64,
    pub standard_error: f64,
}

/// `ym_su3_chain_config` (su3_ensemble.rs:36-46 plus the device options).
#[repr(C)]
#[derive(Clone, Copy, Debug, PartialEq)]
pub struct YmSu3ChainConfig {
    pub extents: [i64; 4],
    pub beta: f64,
    pub master_seed: u64,
    pub chain_id: u64,
    pub thermalization_sweeps: i64,
    pub measurements: i64,
    pub sweeps_between_measurements: i64,
    pub proposal_angle: f64,
    pub autocorrelation_max_lag: i64,
    pub hits: i32,
    pub precision: i32,
    pub hot_start: i32,
    pub reunitarize: i32,
}

/// `ym_su3_chain_report` (su3_ensemble.rs:100-115, measurements separate).
#[repr(C)]
#[derive(Clone, Copy, Debug, Default, PartialEq)]
pub struct YmSu3ChainReport {
    pub derived_seed: u64,
    pub total_sweeps: i64,
    pub proposals: u64,
    pub accepts: u64,
    pub mean_acceptance: f64,
    pub final_unitarity_residual: f64,
    pub final_determinant_residual: f64,
    pub diagnostics: YmChainDiagnostics,
}

/// `ym_su2_chain_config` (su2_ensemble.rs:56-68 plus the device options).
#[repr(C)]
#[derive(Clone, Copy, Debug, PartialEq)]
pub struct YmSu2ChainConfig {
    pub extents: [i64; 4],
    pub beta: f64,
    pub master_seed: u64,
    pub chain_id: u64,
    pub thermalization_sweeps: i64,
    pub measurements: i64,
    pub sweeps_between_measurements: i64,
    pub overrelaxation_sweeps_per_heatbath: i64,
    pub max_heatbath_attempts_per_link: i64,
    pub autocorrelation_max_lag: i64,
    pub precision: i32,
    pub hot_start: i32,
}

/// `ym_su2_chain_report` (su2_ensemble.rs:173-186, measurements separate).
#[repr(C)]
#[derive(Clone, Copy, Debug, Default, PartialEq)]
pub struct YmSu2ChainReport {
    pub derived_seed: u64,
    pub total_heatbath_sweeps: i64,
    pub total_overrelaxation_sweeps: i64,
    pub heatbath_proposals: u64,
    pub heatbath_updates: u64,
    pub mean_heatbath_acceptance: f64,
    pub final_link_unit_residual: f64,
    pub diagnostics: YmChainDiagnostics,
}

extern "C" {
    pub fn ym_lattice_create(group: i32, extents: *const i64, precision: i32, out: *mut *mut YmLattice) -> i32;
    pub fn ym_lattice_destroy(lat: *mut YmLattice) -> i32;
    pub fn ym_lattice_set_identity(lat: *mut YmLattice) -> i32;
    pub fn ym_lattice_hot_start(lat: *mut YmLattice, seed: u64) -> i32;
    pub fn ym_lattice_upload(lat: *mut YmLattice, aos: *const f64, count: i64) -> i32;
    pub fn ym_lattice_download(lat: *const YmLattice, aos: *mut f64, count: i64) -> i32;

    pub fn ym_su3_metropolis_sweeps(lat: *mut YmLattice, beta: f64, proposal_angle: f64, hits: i32, key: u64,
                                    first_sweep: u64, n_sweeps: i32, reunitarize: i32, proposals: *mut u64,
                                    accepts: *mut u64) -> i32;
    pub fn ym_su2_heatbath_sweeps(lat: *mut YmLattice, beta: f64, key: u64, first_sweep: u64, n_sweeps: i32,
                                  max_attempts: i32, proposals: *mut u64, updates: *mut u64) -> i32;
    pub fn ym_su2_overrelax_sweeps(lat: *mut YmLattice, n_sweeps: i32) -> i32;
    pub fn ym_su2_markov_sweeps(lat: *mut YmLattice, beta: f64, key: u64, first_sweep: u64, n_sweeps: i32,
                                overrelaxation_sweeps_per_heatbath: i32, max_attempts: i32, proposals: *mut u64,
                                updates: *mut u64) -> i32;

    pub fn ym_average_plaquette(lat: *const YmLattice, out: *mut f64) -> i32;
    pub fn ym_link_residuals(lat: *const YmLattice, unit: *mut f64, det: *mut f64) -> i32;

    pub fn ym_su3_run_chain(config: *const YmSu3ChainConfig, measurements_out: *mut f64,
                            report: *mut YmSu3ChainReport) -> i32;
    pub fn ym_su2_run_chain(config: *const YmSu2ChainConfig, measurements_out: *mut f64,
                            report: *mut YmSu2ChainReport) -> i32;
    pub fn ym_diagnostics(values: *const f64, n: i64, max_lag: i64, out: *mut YmChainDiagnostics,
                          autocorr: *mut f64) -> i32;

    pub fn ym_derive_chain_seed(master_seed: u64, chain_id: u64) -> u64;
    pub fn ym_derive_su3_chain_seed(master_seed: u64, chain_id: u64) -> u64;
    pub fn ym_status_string(status: i32) -> *const c_char;
}
