//! MI355X drop-in for the hot path of the `yang-mills-reference` crate: the
//! SU(3) Metropolis and SU(2) heat-bath / over-relaxation sweeps, the
//! average plaquette and the chain drivers, on links resident in device
//! memory, through the C ABI of libym_amd.so (include/ym_amd.h).
//!
//! Module and function names follow the reference (`su3_ensemble::
//! metropolis_sweep`, `su2_ensemble::heatbath_sweep`, `run_chain`, ...), and
//! so do argument meaning and error variants.  What differs is the random
//! stream: the device draws counter-based Philox4x32-10 numbers keyed by the
//! chain's derived seed, so the sequential `SplitMix64` argument becomes a
//! [`DeviceStream`] (key, next sweep index), and lattices are
//! [`DeviceLattice`] handles that upload from and download to the
//! reference's link layout.
//!
//! Not compiled in this repository (its build image has no Rust toolchain);
//! see Cargo.toml.

pub mod ffi;
mod lattice;
pub mod su2_ensemble;
pub mod su3_ensemble;

pub use lattice::{diagnostics, ChainDiagnostics, DeviceLattice, DeviceStream, Group, Precision, Status};
