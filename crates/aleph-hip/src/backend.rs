//! `impl Backend` over the C ABI of libaleph_hip.

use std::collections::hash_map::RandomState;
use std::ffi::CStr;
use std::hash::{BuildHasher, Hasher};

use aleph_backend::{Backend, BackendError, MeasurementRecord};
use aleph_core::gate::Param;
use aleph_core::{Complex, Gate, GateInstance, Pauli, PauliString};
use aleph_ir::{Circuit, DiagonalPhase, Instruction, TiledBlock};

use crate::ffi;

/// Largest FP64 state: 2^32 amplitudes of 16 B = 64 GiB of the 288 GB HBM.
pub const MAX_HIP_QUBITS: u32 = 32;
/// Largest FP32 state: 2^33 amplitudes of 8 B.
pub const MAX_HIP_QUBITS_F32: u32 = 33;

/// Whether a HIP device is visible (the `cuda_available` probe of the CLI).
pub fn device_available() -> bool {
    let mut n = 0i32;
    unsafe { ffi::aleph_hip_device_count(&mut n) == ffi::ALEPH_HIP_OK && n > 0 }
}

/// Names the library reports in `aleph_hip_error::kind`, interned so they can
/// fill the reference's `&'static str` error fields.
const KINDS: &[&str] = &[
    "H", "X", "Y", "Z", "S", "Sdg", "T", "Tdg", "Rx", "Ry", "Rz", "Phase", "U3", "Cnot", "Cz", "Swap", "Iswap",
    "IswapDg", "CRx", "CRy", "CRz", "Toffoli", "Ccz", "Unitary1q", "Unitary2q", "UnitaryKq", "Unitary1qDiag",
    "Unitary2qDiag", "diagonal_phase", "unpermute_state", "reset", "measure", "sample", "expectation",
    "probabilities", "allocate", "qaoa", "argument",
    // InvalidState / InvalidPauliString reasons
    "non-finite amplitude norm²", "state norm² deviates from 1 beyond drift budget", "non-finite coefficient",
    "non-finite amplitude",
];

fn intern(raw: &[std::os::raw::c_char]) -> &'static str {
    let bytes: Vec<u8> = raw.iter().take_while(|&&c| c != 0).map(|&c| c as u8).collect();
    let s = String::from_utf8_lossy(&bytes);
    KINDS.iter().copied().find(|k| *k == s).unwrap_or("unknown")
}

fn message(e: &ffi::aleph_hip_error) -> String {
    unsafe { CStr::from_ptr(e.message.as_ptr()) }.to_string_lossy().into_owned()
}

/// `aleph_hip_status` -> `BackendError`, field for field, in the header's
/// enum order (crates/aleph-backend/src/lib.rs:22-81).
/// The `BackendError` of a failed call (status != OK).
pub(crate) fn status_error(ctx: *const ffi::aleph_hip_ctx, status: ffi::aleph_hip_status) -> BackendError {
    match check(ctx, status) {
        Err(e) => e,
        Ok(()) => BackendError::DeviceFailure("status error without a failure".to_string()),
    }
}

fn check(ctx: *const ffi::aleph_hip_ctx, status: ffi::aleph_hip_status) -> Result<(), BackendError> {
    if status == ffi::ALEPH_HIP_OK {
        return Ok(());
    }
    let mut e: ffi::aleph_hip_error = unsafe { std::mem::zeroed() };
    if !ctx.is_null() {
        unsafe { ffi::aleph_hip_last_error(ctx, &mut e) };
    }
    Err(match status {
        ffi::ALEPH_HIP_ERR_QUBIT_OUT_OF_RANGE => {
            BackendError::QubitOutOfRange { qubit: e.qubit, num_qubits: e.num_qubits }
        }
        ffi::ALEPH_HIP_ERR_DUPLICATE_QUBIT => BackendError::DuplicateQubit { qubit: e.qubit },
        ffi::ALEPH_HIP_ERR_ARITY_MISMATCH => BackendError::ArityMismatch {
            kind: intern(&e.kind),
            expected: e.expected as usize,
            got: e.got as usize,
        },
        ffi::ALEPH_HIP_ERR_UNSUPPORTED_GATE => BackendError::UnsupportedGate { kind: intern(&e.kind) },
        ffi::ALEPH_HIP_ERR_UNSUPPORTED_INSTRUCTION => BackendError::UnsupportedInstruction { kind: intern(&e.kind) },
        ffi::ALEPH_HIP_ERR_SYMBOLIC_PARAM => BackendError::SymbolicParam,
        ffi::ALEPH_HIP_ERR_NON_FINITE_PARAM => BackendError::NonFiniteParam { kind: intern(&e.kind) },
        ffi::ALEPH_HIP_ERR_NON_UNITARY_MATRIX => BackendError::NonUnitaryMatrix { deviation: e.value },
        ffi::ALEPH_HIP_ERR_EMPTY_CIRCUIT => BackendError::EmptyCircuit,
        ffi::ALEPH_HIP_ERR_DEGENERATE_MEASUREMENT => {
            BackendError::DegenerateMeasurement { qubit: e.qubit, probability: e.value }
        }
        ffi::ALEPH_HIP_ERR_TOO_MANY_QUBITS => BackendError::TooManyQubits { requested: e.got, limit: e.expected },
        ffi::ALEPH_HIP_ERR_INVALID_PAULI_STRING => BackendError::InvalidPauliString { reason: intern(&e.kind) },
        ffi::ALEPH_HIP_ERR_INVALID_STATE => BackendError::InvalidState { reason: intern(&e.kind) },
        // The IR pipeline is not run inside the library; a planner failure and
        // the device-level codes all carry the library's message.
        ffi::ALEPH_HIP_ERR_OPTIMIZATION => BackendError::DeviceFailure(format!("planner: {}", message(&e))),
        ffi::ALEPH_HIP_ERR_NO_DEVICE => BackendError::DeviceFailure("no HIP device".into()),
        ffi::ALEPH_HIP_ERR_INVALID_ARGUMENT => BackendError::DeviceFailure(format!("invalid argument: {}", message(&e))),
        ffi::ALEPH_HIP_ERR_OUT_OF_MEMORY => BackendError::DeviceFailure(format!("out of memory: {}", message(&e))),
        _ => BackendError::DeviceFailure(message(&e)),
    })
}

fn param(p: &Param) -> Result<f64, BackendError> {
    match p {
        Param::Concrete(v) => Ok(*v),
        Param::Symbolic(_) => Err(BackendError::SymbolicParam),
    }
}

fn push_complex(out: &mut Vec<f64>, z: &Complex) {
    out.push(z.re);
    out.push(z.im);
}

/// A gate as (kind, k, params, interleaved matrix payload), in the header's
/// `Gate` enum order (crates/aleph-core/src/gate/kinds.rs:13-115).
fn encode_gate(g: &Gate) -> Result<(u32, u32, [f64; 3], usize, Vec<f64>), BackendError> {
    use ffi::*;
    let mut params = [0.0; 3];
    let mut payload = Vec::new();
    let (kind, n_params, k) = match g {
        Gate::H => (ALEPH_HIP_GATE_H, 0, 0),
        Gate::X => (ALEPH_HIP_GATE_X, 0, 0),
        Gate::Y => (ALEPH_HIP_GATE_Y, 0, 0),
        Gate::Z => (ALEPH_HIP_GATE_Z, 0, 0),
        Gate::S => (ALEPH_HIP_GATE_S, 0, 0),
        Gate::Sdg => (ALEPH_HIP_GATE_SDG, 0, 0),
        Gate::T => (ALEPH_HIP_GATE_T, 0, 0),
        Gate::Tdg => (ALEPH_HIP_GATE_TDG, 0, 0),
        Gate::Rx(t) => {
            params[0] = param(t)?;
            (ALEPH_HIP_GATE_RX, 1, 0)
        }
        Gate::Ry(t) => {
            params[0] = param(t)?;
            (ALEPH_HIP_GATE_RY, 1, 0)
        }
        Gate::Rz(t) => {
            params[0] = param(t)?;
            (ALEPH_HIP_GATE_RZ, 1, 0)
        }
        Gate::Phase(t) => {
            params[0] = param(t)?;
            (ALEPH_HIP_GATE_PHASE, 1, 0)
        }
        Gate::U3(t, p, l) => {
            params = [param(t)?, param(p)?, param(l)?];
            (ALEPH_HIP_GATE_U3, 3, 0)
        }
        Gate::Cnot => (ALEPH_HIP_GATE_CNOT, 0, 0),
        Gate::Cz => (ALEPH_HIP_GATE_CZ, 0, 0),
        Gate::Swap => (ALEPH_HIP_GATE_SWAP, 0, 0),
        Gate::Iswap => (ALEPH_HIP_GATE_ISWAP, 0, 0),
        Gate::IswapDg => (ALEPH_HIP_GATE_ISWAPDG, 0, 0),
        Gate::CRx(t) => {
            params[0] = param(t)?;
            (ALEPH_HIP_GATE_CRX, 1, 0)
        }
        Gate::CRy(t) => {
            params[0] = param(t)?;
            (ALEPH_HIP_GATE_CRY, 1, 0)
        }
        Gate::CRz(t) => {
            params[0] = param(t)?;
            (ALEPH_HIP_GATE_CRZ, 1, 0)
        }
        Gate::Toffoli => (ALEPH_HIP_GATE_TOFFOLI, 0, 0),
        Gate::Ccz => (ALEPH_HIP_GATE_CCZ, 0, 0),
        Gate::Unitary1q(m) => {
            m.iter().flatten().for_each(|z| push_complex(&mut payload, z));
            (ALEPH_HIP_GATE_UNITARY1Q, 0, 0)
        }
        Gate::Unitary2q(m) => {
            m.iter().flatten().for_each(|z| push_complex(&mut payload, z));
            (ALEPH_HIP_GATE_UNITARY2Q, 0, 0)
        }
        Gate::UnitaryKq { k, data } => {
            data.iter().for_each(|z| push_complex(&mut payload, z));
            (ALEPH_HIP_GATE_UNITARYKQ, 0, u32::from(*k))
        }
        Gate::Unitary1qDiag(d) => {
            d.iter().for_each(|z| push_complex(&mut payload, z));
            (ALEPH_HIP_GATE_UNITARY1QDIAG, 0, 0)
        }
        Gate::Unitary2qDiag(d) => {
            d.iter().for_each(|z| push_complex(&mut payload, z));
            (ALEPH_HIP_GATE_UNITARY2QDIAG, 0, 0)
        }
    };
    Ok((kind, k, params, n_params, payload))
}

fn encode_gate_record(g: &GateInstance, words: &mut Vec<u32>, reals: &mut Vec<f64>) -> Result<(), BackendError> {
    let (kind, k, params, n_params, payload) = encode_gate(&g.gate)?;
    words.extend([kind, g.qubits.len() as u32, g.controls.len() as u32, k]);
    words.extend(g.qubits.iter().copied());
    words.extend(g.controls.iter().copied());
    reals.extend_from_slice(&params[..n_params]);
    reals.extend(payload);
    Ok(())
}

fn encode_diagonal_phase(dp: &DiagonalPhase, words: &mut Vec<u32>, reals: &mut Vec<f64>) {
    words.extend([ffi::ALEPH_HIP_INSTR_DIAGONAL_PHASE, dp.terms.len() as u32]);
    for t in &dp.terms {
        words.push(t.conds.len() as u32);
        for &m in &t.conds {
            words.extend([m as u32, (m >> 32) as u32]);
        }
        reals.push(t.angle);
    }
}

/// The word/real program stream of `aleph_hip_run_program` for a run of
/// instructions without measurements (gates, diagonal phases, tiled blocks,
/// barriers).  `Measure` and `Reset` are rejected here; the circuit driver
/// splits around them.
pub fn encode_program(instructions: &[Instruction]) -> Result<(Vec<u32>, Vec<f64>), BackendError> {
    let mut words = Vec::new();
    let mut reals = Vec::new();
    for inst in instructions {
        match inst {
            Instruction::Gate(g) => encode_gate_record(g, &mut words, &mut reals)?,
            Instruction::DiagonalPhase(dp) => encode_diagonal_phase(dp, &mut words, &mut reals),
            Instruction::TiledBlock(tb) => {
                for g in &tb.gates {
                    encode_gate_record(g, &mut words, &mut reals)?;
                }
            }
            Instruction::Barrier(qs) => {
                words.extend([ffi::ALEPH_HIP_INSTR_BARRIER, qs.len() as u32]);
                words.extend(qs.iter().copied());
            }
            Instruction::Measure { .. } => return Err(BackendError::UnsupportedInstruction { kind: "measure" }),
            Instruction::Reset(_) => return Err(BackendError::UnsupportedInstruction { kind: "reset" }),
        }
    }
    Ok((words, reals))
}

/// Planner options of the whole-circuit path (`aleph_hip_plan_opts`).
#[derive(Clone, Copy, Debug)]
pub struct HipPlanOptions {
    /// `ALEPH_HIP_PLANNER_RULE` (default), `_LEARNED` or `_UNFUSED`.
    pub planner: u32,
    /// Fused passes as kernels generated per pass structure (hiprtc at plan creation).
    pub generated_kernels: bool,
    /// 0, or 3..=6: dense k-qubit blocks instead of LDS passes.
    pub kq_width: u32,
}

impl Default for HipPlanOptions {
    fn default() -> Self {
        Self { planner: ffi::ALEPH_HIP_PLANNER_RULE, generated_kernels: true, kq_width: 0 }
    }
}

/// A device state vector; freed (stream-ordered) on drop.  It keeps the
/// context it was allocated on, which must outlive every call on the state.
pub struct HipState {
    st: *mut ffi::aleph_hip_state,
    ctx: *mut ffi::aleph_hip_ctx,
    num_qubits: u32,
}

unsafe impl Send for HipState {}

impl HipState {
    pub fn num_qubits(&self) -> u32 {
        self.num_qubits
    }

    /// Amplitudes copied to the host as complex128 (FP32 states widen).
    pub fn amplitudes(&self) -> Result<Vec<Complex>, BackendError> {
        let len = 1usize << self.num_qubits;
        let mut raw = vec![0.0f64; 2 * len];
        check(self.ctx, unsafe { ffi::aleph_hip_state_read(self.ctx, self.st, raw.as_mut_ptr()) })?;
        Ok(raw.chunks_exact(2).map(|c| Complex::new(c[0], c[1])).collect())
    }
}

/// A finished state together with the backend that owns its context, read
/// back only when asked: the device side of aleph-py's lazy amplitude export
/// (crates/aleph-py/src/lazy_amplitudes.rs).  Fields drop in order: the state
/// before its context.
pub struct HipSvReadout {
    state: HipState,
    _backend: HipSvBackend,
}

impl HipSvReadout {
    pub fn new(backend: HipSvBackend, state: HipState) -> Self {
        Self { state, _backend: backend }
    }

    pub fn try_amplitudes_vec(&self) -> Result<Vec<Complex>, BackendError> {
        self.state.amplitudes()
    }
}

impl Drop for HipState {
    fn drop(&mut self) {
        if !self.st.is_null() {
            unsafe { ffi::aleph_hip_state_free(self.st) };
        }
    }
}

/// Dense state vector on one MI355X, FP64 amplitudes.
pub struct HipSvBackend {
    ctx: *mut ffi::aleph_hip_ctx,
    precision: u32,
    plan: HipPlanOptions,
}

unsafe impl Send for HipSvBackend {}

fn entropy_seed() -> u64 {
    // std seeds RandomState's keys randomly per process.
    let mut h = RandomState::new().build_hasher();
    h.write_u64(std::process::id() as u64);
    h.finish()
}

impl HipSvBackend {
    /// Device 0, a random seed.
    pub fn new() -> Result<Self, BackendError> {
        Self::open(entropy_seed(), ffi::ALEPH_HIP_F64)
    }

    pub fn with_seed(seed: u64) -> Result<Self, BackendError> {
        Self::open(seed, ffi::ALEPH_HIP_F64)
    }

    fn open(seed: u64, precision: u32) -> Result<Self, BackendError> {
        let mut ctx = std::ptr::null_mut();
        check(std::ptr::null(), unsafe { ffi::aleph_hip_ctx_create(0, seed, &mut ctx) })?;
        Ok(Self { ctx, precision, plan: HipPlanOptions::default() })
    }

    /// Lower the qubit cap (`with_qubit_cap`, crates/aleph-sv/src/backend.rs:46-50).
    pub fn with_qubit_cap(self, cap: u32) -> Self {
        unsafe { ffi::aleph_hip_ctx_set_qubit_cap(self.ctx, cap) };
        self
    }

    pub fn with_plan_options(mut self, plan: HipPlanOptions) -> Self {
        self.plan = plan;
        self
    }

    pub(crate) fn raw_ctx(&self) -> *mut ffi::aleph_hip_ctx {
        self.ctx
    }

    fn opts(&self) -> ffi::aleph_hip_plan_opts {
        ffi::aleph_hip_plan_opts {
            planner: self.plan.planner,
            flags: if self.plan.generated_kernels { ffi::ALEPH_HIP_PLAN_JIT } else { 0 },
            kq_width: self.plan.kq_width,
            fuse_1q: 1,
            ..Default::default()
        }
    }

    /// One measurement-free run of instructions as one fused plan.
    pub fn run_program(&mut self, state: &mut HipState, instructions: &[Instruction]) -> Result<(), BackendError> {
        if instructions.is_empty() {
            return Ok(());
        }
        let (words, reals) = encode_program(instructions)?;
        let opts = self.opts();
        check(self.ctx, unsafe {
            ffi::aleph_hip_run_program(
                self.ctx,
                state.st,
                words.as_ptr(),
                words.len(),
                reals.as_ptr(),
                reals.len(),
                &opts,
                std::ptr::null_mut(),
            )
        })
    }

    /// Whole-circuit execution with the outcome records of `run_with_outcomes`
    /// (crates/aleph-backend/src/lib.rs:211-248): every segment between
    /// measurements is one fused plan.
    pub fn run_circuit_with_outcomes(
        &mut self,
        circuit: &Circuit,
    ) -> Result<(HipState, Vec<MeasurementRecord>), BackendError> {
        if circuit.num_qubits() == 0 && circuit.is_empty() {
            return Err(BackendError::EmptyCircuit);
        }
        let mut state = self.allocate(circuit.num_qubits())?;
        let mut outcomes = Vec::new();
        let insts = circuit.instructions();
        let mut begin = 0;
        for (idx, inst) in insts.iter().enumerate() {
            match inst {
                Instruction::Measure { qubit, clbit } => {
                    self.run_program(&mut state, &insts[begin..idx])?;
                    begin = idx + 1;
                    let outcome = self.measure(&mut state, *qubit)?;
                    outcomes.push(MeasurementRecord { instruction_index: idx, qubit: *qubit, clbit: *clbit, outcome });
                }
                Instruction::Reset(_) => return Err(BackendError::UnsupportedInstruction { kind: "reset" }),
                _ => {}
            }
        }
        self.run_program(&mut state, &insts[begin..])?;
        Ok((state, outcomes))
    }

    /// Sum of Pauli-string expectations in one call (`expectation_pauli_sum`,
    /// crates/aleph-backend/src/lib.rs:422-432): strings sharing an X mask are
    /// evaluated together, up to 16 per read of the state.
    pub fn expectation_pauli_sum(&mut self, state: &HipState, paulis: &[PauliString]) -> Result<f64, BackendError> {
        let mut coeffs = Vec::with_capacity(paulis.len());
        let mut offsets = Vec::with_capacity(paulis.len() + 1);
        let mut qubits = Vec::new();
        let mut codes = Vec::new();
        offsets.push(0u32);
        for p in paulis {
            coeffs.push(p.coefficient);
            for (q, op) in &p.terms {
                qubits.push(*q);
                codes.push(pauli_code(*op));
            }
            offsets.push(qubits.len() as u32);
        }
        let mut out = 0.0;
        check(self.ctx, unsafe {
            ffi::aleph_hip_expectation_sum(
                self.ctx,
                state.st,
                paulis.len() as u32,
                coeffs.as_ptr(),
                offsets.as_ptr(),
                qubits.as_ptr(),
                codes.as_ptr(),
                &mut out,
            )
        })?;
        Ok(out)
    }

    /// Linear XEB of the exact state, 2^n sum p^2 - 1 (benches/src/lib.rs:29).
    pub fn linear_xeb(&mut self, state: &HipState) -> Result<f64, BackendError> {
        let mut out = 0.0;
        check(self.ctx, unsafe { ffi::aleph_hip_linear_xeb(self.ctx, state.st, &mut out) })?;
        Ok(out)
    }

}

impl Drop for HipSvBackend {
    fn drop(&mut self) {
        if !self.ctx.is_null() {
            unsafe { ffi::aleph_hip_ctx_destroy(self.ctx) };
        }
    }
}

fn pauli_code(p: Pauli) -> u8 {
    match p {
        Pauli::I => 0,
        Pauli::X => 1,
        Pauli::Y => 2,
        Pauli::Z => 3,
    }
}

impl Backend for HipSvBackend {
    type State = HipState;

    fn allocate(&mut self, num_qubits: u32) -> Result<HipState, BackendError> {
        let mut st = std::ptr::null_mut();
        check(self.ctx, unsafe { ffi::aleph_hip_state_alloc(self.ctx, num_qubits, self.precision, &mut st) })?;
        Ok(HipState { st, ctx: self.ctx, num_qubits })
    }

    fn apply_gate(&mut self, state: &mut HipState, gate: &GateInstance) -> Result<(), BackendError> {
        let (kind, k, params, _, payload) = encode_gate(&gate.gate)?;
        let g = ffi::aleph_hip_gate {
            kind,
            n_qubits: gate.qubits.len() as u32,
            n_controls: gate.controls.len() as u32,
            k,
            qubits: gate.qubits.as_ptr(),
            controls: gate.controls.as_ptr(),
            params,
            matrix: if payload.is_empty() { std::ptr::null() } else { payload.as_ptr() },
        };
        check(self.ctx, unsafe { ffi::aleph_hip_apply_gate(self.ctx, state.st, &g) })
    }

    fn measure(&mut self, state: &mut HipState, qubit: u32) -> Result<bool, BackendError> {
        let mut outcome = 0i32;
        check(self.ctx, unsafe { ffi::aleph_hip_measure(self.ctx, state.st, qubit, &mut outcome) })?;
        Ok(outcome != 0)
    }

    fn sample(&mut self, state: &HipState, shots: u32) -> Result<Vec<u64>, BackendError> {
        let mut out = vec![0u64; shots as usize];
        check(self.ctx, unsafe { ffi::aleph_hip_sample(self.ctx, state.st, shots, out.as_mut_ptr()) })?;
        Ok(out)
    }

    fn expectation_value(&mut self, state: &HipState, pauli: &PauliString) -> Result<f64, BackendError> {
        let qubits: Vec<u32> = pauli.terms.iter().map(|(q, _)| *q).collect();
        let codes: Vec<u8> = pauli.terms.iter().map(|(_, p)| pauli_code(*p)).collect();
        let mut out = 0.0;
        check(self.ctx, unsafe {
            ffi::aleph_hip_expectation(
                self.ctx,
                state.st,
                pauli.coefficient,
                qubits.as_ptr(),
                codes.as_ptr(),
                qubits.len() as u32,
                &mut out,
            )
        })?;
        Ok(out)
    }

    fn probabilities(&mut self, state: &HipState, qubits: &[u32]) -> Result<Vec<f64>, BackendError> {
        if qubits.len() >= 64 {
            return Err(BackendError::QubitOutOfRange { qubit: qubits[63], num_qubits: state.num_qubits });
        }
        let mut out = vec![0.0; 1usize << qubits.len()];
        check(self.ctx, unsafe {
            ffi::aleph_hip_probabilities(self.ctx, state.st, qubits.as_ptr(), qubits.len() as u32, out.as_mut_ptr())
        })?;
        Ok(out)
    }

    fn apply_diagonal_phase(&mut self, state: &mut HipState, dp: &DiagonalPhase) -> Result<(), BackendError> {
        let mut angles = Vec::with_capacity(dp.terms.len());
        let mut conds = Vec::new();
        let mut offsets = vec![0u32];
        for t in &dp.terms {
            angles.push(t.angle);
            conds.extend(t.conds.iter().copied());
            offsets.push(conds.len() as u32);
        }
        check(self.ctx, unsafe {
            ffi::aleph_hip_apply_diagonal_phase(
                self.ctx,
                state.st,
                angles.len() as u32,
                angles.as_ptr(),
                conds.as_ptr(),
                offsets.as_ptr(),
            )
        })
    }

    /// The tiled block's gates as one fused plan (its LDS passes are the
    /// device analogue of the CPU's tile-major replay).
    fn apply_tiled_block(&mut self, state: &mut HipState, block: &TiledBlock) -> Result<(), BackendError> {
        let insts: Vec<Instruction> = block.gates.iter().cloned().map(Instruction::Gate).collect();
        self.run_program(state, &insts)
    }

    fn unpermute_state(&mut self, state: &mut HipState, perm: &[u32]) -> Result<(), BackendError> {
        if perm.len() != state.num_qubits as usize {
            return Err(BackendError::InvalidState { reason: "unknown" });
        }
        check(self.ctx, unsafe { ffi::aleph_hip_unpermute(self.ctx, state.st, perm.as_ptr()) })
    }
}

/// Dense state vector on one MI355X, FP32 amplitudes with FP64 readout
/// accumulation (`CudaSvBackendF32`'s contract, crates/aleph-cuda/src/sv/fp32.rs).
pub struct HipSvBackendF32(HipSvBackend);

impl HipSvBackendF32 {
    /// The inner backend, e.g. to hand to [`HipSvReadout::new`].
    pub fn into_inner(self) -> HipSvBackend {
        self.0
    }

    pub fn new() -> Result<Self, BackendError> {
        HipSvBackend::open(entropy_seed(), ffi::ALEPH_HIP_F32).map(Self)
    }

    pub fn with_seed(seed: u64) -> Result<Self, BackendError> {
        HipSvBackend::open(seed, ffi::ALEPH_HIP_F32).map(Self)
    }

    pub fn with_qubit_cap(self, cap: u32) -> Self {
        Self(self.0.with_qubit_cap(cap))
    }

    pub fn with_plan_options(self, plan: HipPlanOptions) -> Self {
        Self(self.0.with_plan_options(plan))
    }
}

impl std::ops::Deref for HipSvBackendF32 {
    type Target = HipSvBackend;
    fn deref(&self) -> &HipSvBackend {
        &self.0
    }
}

impl std::ops::DerefMut for HipSvBackendF32 {
    fn deref_mut(&mut self) -> &mut HipSvBackend {
        &mut self.0
    }
}

impl Backend for HipSvBackendF32 {
    type State = HipState;

    fn allocate(&mut self, num_qubits: u32) -> Result<HipState, BackendError> {
        self.0.allocate(num_qubits)
    }
    fn apply_gate(&mut self, state: &mut HipState, gate: &GateInstance) -> Result<(), BackendError> {
        self.0.apply_gate(state, gate)
    }
    fn measure(&mut self, state: &mut HipState, qubit: u32) -> Result<bool, BackendError> {
        self.0.measure(state, qubit)
    }
    fn sample(&mut self, state: &HipState, shots: u32) -> Result<Vec<u64>, BackendError> {
        self.0.sample(state, shots)
    }
    fn expectation_value(&mut self, state: &HipState, pauli: &PauliString) -> Result<f64, BackendError> {
        self.0.expectation_value(state, pauli)
    }
    fn probabilities(&mut self, state: &HipState, qubits: &[u32]) -> Result<Vec<f64>, BackendError> {
        self.0.probabilities(state, qubits)
    }
    fn apply_diagonal_phase(&mut self, state: &mut HipState, dp: &DiagonalPhase) -> Result<(), BackendError> {
        self.0.apply_diagonal_phase(state, dp)
    }
    fn apply_tiled_block(&mut self, state: &mut HipState, block: &TiledBlock) -> Result<(), BackendError> {
        self.0.apply_tiled_block(state, block)
    }
    fn unpermute_state(&mut self, state: &mut HipState, perm: &[u32]) -> Result<(), BackendError> {
        self.0.unpermute_state(state, perm)
    }
}
