//! MI355X (gfx950) dense state-vector backend for aleph.
//!
//! The compute lives in `libaleph_hip.so` (hand-written HIP kernels and a C++
//! host runtime); this crate binds its C ABI (`include/aleph_hip.h`) and
//! implements [`aleph_backend::Backend`] on top of it, so `--backend hip` /
//! `backend="hip"` sit beside `sv` and `cuda`:
//!
//! * [`HipSvBackend`] (FP64 amplitudes) and [`HipSvBackendF32`] (FP32
//!   amplitudes, FP64 readout accumulation) implement `Backend` gate by gate,
//!   like `CudaSvBackend` / `CudaSvBackendF32` (crates/aleph-cuda/src/sv);
//! * [`HipSvBackend::run_circuit_with_outcomes`] is the whole-circuit path,
//!   the analogue of `prepare_cuda_execution` + `CudaExecutionPlan::execute`
//!   (crates/aleph-cuda/src/execution.rs:78-273): each measurement-free
//!   segment of the circuit is planned into fused LDS passes and executed in
//!   one library call, measurements in between go through `Backend::measure`,
//!   with the outcome records of `run_with_outcomes`
//!   (crates/aleph-backend/src/lib.rs:211-248).
//!
//! Everything is gated on `cfg(all(target_os = "linux", feature = "hip"))`;
//! without the feature the crate is empty, as `aleph-cuda` is without `cuda`.

#[cfg(all(target_os = "linux", feature = "hip"))]
pub mod ffi;

#[cfg(all(target_os = "linux", feature = "hip"))]
mod backend;

/// `run_noisy` / `run_noisy_sparse` on the device (feature `noise`).
#[cfg(all(target_os = "linux", feature = "noise"))]
pub mod noise;

#[cfg(all(target_os = "linux", feature = "hip"))]
pub use backend::{
    device_available, encode_program, HipPlanOptions, HipState, HipSvBackend, HipSvBackendF32, HipSvReadout,
    MAX_HIP_QUBITS, MAX_HIP_QUBITS_F32,
};
