//! Noisy execution on the device: `run_noisy` / `run_noisy_sparse` of
//! crates/aleph-sv/src/noise/execution.rs:360-400 over the library's
//! trajectory kernels (`aleph_hip_run_noisy`).  The caller's `NoiseModel`
//! is resolved here exactly as the reference resolves it
//! (`NoiseModel::errors_for`, all-qubit attachments first), then every
//! trajectory runs in one workgroup on the GPU.  Each shot draws from its own
//! splitmix64 stream keyed by the reference's shot seed, so outcomes depend
//! only on (seed, shot); they are not bit-identical to the CPU driver's
//! ChaCha12 draws, only equal in distribution.

use aleph_backend::BackendError;
use aleph_core::Pauli;
use aleph_ir::{Circuit, Instruction};
use aleph_sv::noise::{Counts, NoiseError, NoiseModel, QuantumError, SparseCounts};

use crate::backend::{encode_program, HipSvBackend};
use crate::ffi;

fn pauli_code(p: Pauli) -> u8 {
    match p {
        Pauli::I => 0,
        Pauli::X => 1,
        Pauli::Y => 2,
        Pauli::Z => 3,
    }
}

/// Channel payloads kept alive while the library reads them.
struct Channels {
    records: Vec<ffi::aleph_hip_channel>,
    _weights: Vec<Vec<f64>>,
    _codes: Vec<Vec<u8>>,
    _kraus: Vec<Vec<f64>>,
}

fn resolve(circuit: &Circuit, model: &NoiseModel) -> Result<(Vec<Instruction>, Channels), NoiseError> {
    let mut program = Vec::new();
    let mut records = Vec::new();
    let (mut weights, mut codes, mut kraus) = (Vec::new(), Vec::new(), Vec::new());
    for inst in circuit.instructions() {
        match inst {
            Instruction::Gate(gi) => {
                let after = program.len() as u32;
                for err in model.errors_for(gi.gate.name(), &gi.qubits) {
                    if err.arity() != gi.qubits.len() || gi.qubits.len() > 2 {
                        return Err(NoiseError::InvalidChannel { reason: "attachment arity does not match gate" });
                    }
                    let mut q = [0u32; 2];
                    for (j, &x) in gi.qubits.iter().enumerate() {
                        q[j] = x;
                    }
                    let mut rec = ffi::aleph_hip_channel {
                        after,
                        kind: 0,
                        n_qubits: gi.qubits.len() as u32,
                        n_branches: 0,
                        qubits: q,
                        weights: std::ptr::null(),
                        paulis: std::ptr::null(),
                        kraus: std::ptr::null(),
                    };
                    match err {
                        QuantumError::Pauli(c) => {
                            let w: Vec<f64> = c.terms.iter().map(|(p, _)| *p).collect();
                            let cs: Vec<u8> =
                                c.terms.iter().flat_map(|(_, ps)| ps.iter().map(|&p| pauli_code(p))).collect();
                            rec.kind = ffi::ALEPH_HIP_CHANNEL_PAULI;
                            rec.n_branches = c.terms.len() as u32;
                            rec.weights = w.as_ptr();
                            rec.paulis = cs.as_ptr();
                            weights.push(w);
                            codes.push(cs);
                        }
                        QuantumError::Kraus(c) => {
                            let mut k = Vec::with_capacity(8 * c.kraus.len());
                            for m in &c.kraus {
                                for row in m {
                                    for z in row {
                                        k.extend([z.re, z.im]);
                                    }
                                }
                            }
                            rec.kind = ffi::ALEPH_HIP_CHANNEL_KRAUS;
                            rec.n_branches = c.kraus.len() as u32;
                            rec.kraus = k.as_ptr();
                            kraus.push(k);
                        }
                    }
                    records.push(rec);
                }
                program.push(inst.clone());
            }
            Instruction::DiagonalPhase(_) | Instruction::Barrier(_) => program.push(inst.clone()),
            Instruction::Measure { .. } => return Err(NoiseError::MidCircuit { kind: "measure" }),
            Instruction::Reset(_) => return Err(NoiseError::MidCircuit { kind: "reset" }),
            Instruction::TiledBlock(_) => return Err(NoiseError::Unsupported { kind: "tiled-block" }),
        }
    }
    Ok((program, Channels { records, _weights: weights, _codes: codes, _kraus: kraus }))
}

/// The library's InvalidChannel reasons, as the reference's static strings.
fn channel_reason(msg: &str) -> &'static str {
    const REASONS: &[&str] = &[
        "invalid Pauli probability or term arity",
        "Pauli probabilities must sum to one",
        "non-finite Kraus matrix",
        "Kraus operators must be trace preserving",
        "invalid readout probability or qubit",
    ];
    REASONS.iter().copied().find(|r| msg.contains(r)).unwrap_or("invalid noise channel")
}

/// Outcome (basis index) of every shot, shot i being absolute shot `shot_offset + i`.
pub fn outcomes(
    backend: &mut HipSvBackend,
    circuit: &Circuit,
    model: &NoiseModel,
    shots: u32,
    seed: u64,
    shot_offset: u64,
) -> Result<Vec<u64>, NoiseError> {
    let n = circuit.num_qubits();
    let (program, channels) = resolve(circuit, model)?;
    let (words, reals) = encode_program(&program)?;
    let mut readout = vec![0.0f64; 4 * n.max(1) as usize];
    let mut mask = 0u64;
    for q in 0..n.min(64) {
        if let Some(ro) = model.readout_error(q) {
            readout[4 * q as usize..4 * q as usize + 4].copy_from_slice(&[ro.m[0][0], ro.m[0][1], ro.m[1][0], ro.m[1][1]]);
            mask |= 1u64 << q;
        }
    }
    let mut out = vec![0u64; shots as usize];
    let ctx = backend.raw_ctx();
    let status = unsafe {
        ffi::aleph_hip_run_noisy(
            ctx,
            n,
            words.as_ptr(),
            words.len(),
            reals.as_ptr(),
            reals.len(),
            channels.records.as_ptr(),
            channels.records.len() as u32,
            readout.as_ptr(),
            mask,
            shots,
            seed,
            shot_offset,
            0,
            out.as_mut_ptr(),
        )
    };
    if status != ffi::ALEPH_HIP_OK {
        let mut e: ffi::aleph_hip_error = unsafe { std::mem::zeroed() };
        unsafe { ffi::aleph_hip_last_error(ctx, &mut e) };
        let kind = unsafe { std::ffi::CStr::from_ptr(e.kind.as_ptr()) }.to_string_lossy().into_owned();
        let msg = unsafe { std::ffi::CStr::from_ptr(e.message.as_ptr()) }.to_string_lossy().into_owned();
        return Err(match kind.as_str() {
            "invalid-channel" => NoiseError::InvalidChannel { reason: channel_reason(&msg) },
            "unsupported" => NoiseError::Unsupported { kind: "dense gate on more than 3 qubits" },
            _ => NoiseError::Backend(crate::backend::status_error(ctx, status)),
        });
    }
    Ok(out)
}

/// Dense histogram of length 2^n (`run_noisy`, noise/mod.rs:80-90).
pub fn run_noisy(
    backend: &mut HipSvBackend,
    circuit: &Circuit,
    model: &NoiseModel,
    shots: u32,
    seed: u64,
) -> Result<Counts, NoiseError> {
    let n = circuit.num_qubits();
    if n > 24 {
        return Err(NoiseError::Backend(BackendError::TooManyQubits { requested: n, limit: 24 }));
    }
    let mut hist = vec![0u64; 1usize << n];
    for idx in outcomes(backend, circuit, model, shots, seed, 0)? {
        hist[idx as usize] += 1;
    }
    Ok(hist)
}

/// Ascending (basis index, count) pairs (`run_noisy_sparse`, execution.rs:378-400).
pub fn run_noisy_sparse(
    backend: &mut HipSvBackend,
    circuit: &Circuit,
    model: &NoiseModel,
    shots: u32,
    seed: u64,
) -> Result<SparseCounts, NoiseError> {
    let mut idx = outcomes(backend, circuit, model, shots, seed, 0)?;
    idx.sort_unstable();
    let mut counts: SparseCounts = Vec::new();
    for i in idx {
        match counts.last_mut() {
            Some((last, c)) if *last == i => *c += 1,
            _ => counts.push((i, 1)),
        }
    }
    Ok(counts)
}
