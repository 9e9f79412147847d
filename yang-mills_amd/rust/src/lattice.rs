//! Device lattice handle, status, and the counter-based stream that stands in
//! for the reference's sequential `SplitMix64` (su2_ensemble.rs:24-45).

use std::ffi::CStr;
use std::fmt;

use crate::ffi;

/// Gauge group of a lattice (`group` argument of `ym_lattice_create`).
#[derive(Clone, Copy, Debug, PartialEq, Eq)]
pub enum Group {
    Su2 = 2,
    Su3 = 3,
}

/// Link storage precision on the device (`precision` of `ym_lattice_create`).
#[derive(Clone, Copy, Debug, PartialEq, Eq)]
pub enum Precision {
    Fp32 = 32,
    Fp64 = 64,
}

/// A `YM_*` status other than `YM_OK` (include/ym_amd.h lists each with the
/// reference error variant it stands for).
#[derive(Clone, Copy, Debug, PartialEq, Eq)]
pub struct Status(pub i32);

pub(crate) const YM_INVALID_ARGUMENT: i32 = 96;

impl Status {
    pub fn check(code: i32) -> Result<(), Status> {
        if code == 0 {
            Ok(())
        } else {
            Err(Status(code))
        }
    }

    pub fn message(self) -> String {
        // SAFETY: ym_status_string returns a static NUL-terminated string.
        let text = unsafe { ffi::ym_status_string(self.0) };
        if text.is_null() {
            return format!("status {}", self.0);
        }
        unsafe { CStr::from_ptr(text) }.to_string_lossy().into_owned()
    }
}

impl fmt::Display for Status {
    fn fmt(&self, f: &mut fmt::Formatter<'_>) -> fmt::Result {
        write!(f, "{} (status {})", self.message(), self.0)
    }
}

impl std::error::Error for Status {}

/// The random stream of one chain.  The device draws
/// Philox4x32-10(key; site, mu, hit, sweep), so instead of a sequential
/// state the stream is the chain key (the reference's derived seed) and the
/// index of the next sweep; a sweep call advances `next_sweep`.  Saving and
/// restoring this pair resumes a chain bit for bit.
#[derive(Clone, Copy, Debug, PartialEq, Eq)]
pub struct DeviceStream {
    pub key: u64,
    pub next_sweep: u64,
}

impl DeviceStream {
    /// Key derive_su3_chain_seed(master, chain) (su3_ensemble.rs:189-192).
    pub fn su3(master_seed: u64, chain_id: u64) -> Self {
        let key = unsafe { ffi::ym_derive_su3_chain_seed(master_seed, chain_id) };
        Self { key, next_sweep: 0 }
    }

    /// Key derive_chain_seed(master, chain) (su2_ensemble.rs:52-55).
    pub fn su2(master_seed: u64, chain_id: u64) -> Self {
        let key = unsafe { ffi::ym_derive_chain_seed(master_seed, chain_id) };
        Self { key, next_sweep: 0 }
    }
}

/// Links resident in device memory: one `ym_lattice` handle, with its own
/// HIP stream.  A handle is used from one host thread at a time.
pub struct DeviceLattice {
    handle: *mut ffi::YmLattice,
    group: Group,
    precision: Precision,
    extents: [usize; 4],
}

impl DeviceLattice {
    /// Lattice4::new_identity / Su3Lattice4::new_identity (cold start).
    pub fn new_identity(group: Group, extents: [usize; 4], precision: Precision) -> Result<Self, Status> {
        let ext: [i64; 4] = extents.map(|e| e as i64);
        let mut handle = std::ptr::null_mut();
        Status::check(unsafe { ffi::ym_lattice_create(group as i32, ext.as_ptr(), precision as i32, &mut handle) })?;
        Ok(Self { handle, group, precision, extents })
    }

    /// Haar-random links from Philox keyed by `seed` (no reference
    /// counterpart: the reference starts cold).
    pub fn hot_start(&mut self, seed: u64) -> Result<(), Status> {
        Status::check(unsafe { ffi::ym_lattice_hot_start(self.handle, seed) })
    }

    /// Every link from a host image in the reference's layout: link index
    /// site * 4 + mu, sites lexicographic with x3 fastest, 4 f64 per SU(2)
    /// link (a0..a3) or 18 per SU(3) link (row-major re, im).  Fails closed,
    /// storing nothing, on a link the reference's types cannot hold (the
    /// checks of Su3::from_checked_matrix and of the SU(2) constructors).
    pub fn upload(&mut self, links: &[f64]) -> Result<(), Status> {
        Status::check(unsafe { ffi::ym_lattice_upload(self.handle, links.as_ptr(), links.len() as i64) })
    }

    /// Every link as a host image in the reference's layout.
    pub fn download(&self) -> Result<Vec<f64>, Status> {
        let count = self.link_count() * self.reals_per_link();
        let mut links = vec![0.0; count];
        Status::check(unsafe { ffi::ym_lattice_download(self.handle, links.as_mut_ptr(), count as i64) })?;
        Ok(links)
    }

    /// (maximum unitarity residual, maximum determinant residual) over links.
    pub fn link_residuals(&self) -> Result<(f64, f64), Status> {
        let (mut unit, mut det) = (0.0, 0.0);
        Status::check(unsafe { ffi::ym_link_residuals(self.handle, &mut unit, &mut det) })?;
        Ok((unit, det))
    }

    pub fn extents(&self) -> [usize; 4] {
        self.extents
    }

    pub fn group(&self) -> Group {
        self.group
    }

    pub fn precision(&self) -> Precision {
        self.precision
    }

    pub fn link_count(&self) -> usize {
        self.extents.iter().product::<usize>() * 4
    }

    fn reals_per_link(&self) -> usize {
        match self.group {
            Group::Su2 => 4,
            Group::Su3 => 18,
        }
    }

    pub(crate) fn raw(&self) -> *mut ffi::YmLattice {
        self.handle
    }
}

impl Drop for DeviceLattice {
    fn drop(&mut self) {
        // SAFETY: the handle came from ym_lattice_create and is dropped once.
        unsafe {
            ffi::ym_lattice_destroy(self.handle);
        }
    }
}

/// `ChainDiagnostics` (su2_ensemble.rs:162-170): the device report plus the
/// autocorrelation vector, recomputed on the host by `ym_diagnostics`.
#[derive(Clone, Debug, PartialEq)]
pub struct ChainDiagnostics {
    pub mean: f64,
    pub variance: f64,
    pub autocorrelation: Vec<f64>,
    pub positive_window_lag: usize,
    pub integrated_autocorrelation_time: f64,
    pub effective_sample_size: f64,
    pub standard_error: f64,
}

/// su2_ensemble.rs:375-463 diagnostics on the host side of the library.
pub fn diagnostics(values: &[f64], max_lag: usize) -> Result<ChainDiagnostics, Status> {
    let mut out = ffi::YmChainDiagnostics::default();
    let mut autocorrelation = vec![0.0; max_lag + 1];
    Status::check(unsafe {
        ffi::ym_diagnostics(values.as_ptr(), values.len() as i64, max_lag as i64, &mut out,
                            autocorrelation.as_mut_ptr())
    })?;
    Ok(ChainDiagnostics {
        mean: out.mean,
        variance: out.variance,
        autocorrelation,
        positive_window_lag: out.positive_window_lag as usize,
        integrated_autocorrelation_time: out.tau_int,
        effective_sample_size: out.ess,
        standard_error: out.standard_error,
    })
}
