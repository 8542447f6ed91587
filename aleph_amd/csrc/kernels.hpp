// Host-callable launchers of the gfx950 kernels.  T is double or float
// (complex128 / complex64 amplitudes); every launch is ordered on `stream`.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "plan.hpp"

namespace aleph {

// ---- state preparation and per-gate kernels (k_gates.hip) ----
template <typename T> void launch_init_basis(hipStream_t s, void* state, uint32_t n);
template <typename T>
void launch_1q(hipStream_t s, void* state, uint32_t n, uint32_t target, uint64_t cmask, const double* m8);
template <typename T>
void launch_2q(hipStream_t s, void* state, uint32_t n, uint32_t q_msb, uint32_t q_lsb, uint64_t cmask,
               const double* m32);
// Dense 2^k x 2^k (k <= 6) on qubits (MSB first); `dmat` is a device copy of
// the interleaved matrix in the state's precision, row-major.
template <typename T>
void launch_kq(hipStream_t s, void* state, uint32_t n, uint32_t k, const uint32_t* qubits, uint64_t cmask,
               const void* dmat);
// Diagonal on k qubits (MSB first); `ddiag` device copy of 2^k complex in double.
// The same for k in [4, 6] on the matrix cores (k_kq_mfma.hip); launch_kq
// routes those widths here.
template <typename T>
void launch_kq_mfma(hipStream_t s, void* state, uint32_t n, uint32_t k, const uint32_t* qubits, uint64_t cmask,
                    const void* dmat);
template <typename T>
void launch_diag(hipStream_t s, void* state, uint32_t n, uint32_t k, const uint32_t* qubits, uint64_t cmask,
                 const double* ddiag);
template <typename T>
void launch_swap(hipStream_t s, void* state, uint32_t n, uint32_t a, uint32_t b, uint64_t cmask);
// Parity phase polynomial (DiagonalPhase semantics), CSR on the device.
template <typename T>
void launch_phase_poly(hipStream_t s, void* state, uint32_t n, uint32_t n_terms, const double* angles,
                       const uint64_t* conds, const uint32_t* offsets);
template <typename T>
void launch_permute(hipStream_t s, const void* src, void* dst, uint32_t n, const uint32_t* perm);
void launch_widen_f32(hipStream_t s, const void* src_f32, void* dst_f64, uint64_t len);

// ---- qubit permutations that are involutions, in place in one sweep (k_exchange.hip) ----
// sigma[q] = the qubit position bit q moves to; sigma o sigma = id.  Before permuting, the
// n_gates 2x2 matrices (8 doubles each, row-major interleaved) act in order on qubits
// gate_qubits[], which must lie among the low min(5, n) qubits and their images.
constexpr uint32_t kPermMaxGates = 8;
template <typename T>
void launch_bitperm(hipStream_t s, void* state, uint32_t n, const uint32_t* sigma, uint32_t n_gates = 0,
                    const uint32_t* gate_qubits = nullptr, const double* gate_mats = nullptr);

// ---- rank-slice exchange of a distributed state on one device (k_exchange.hip) ----
constexpr uint32_t kMaxExchangeBits = 16;
struct ExchangeBits {
  uint32_t rank_bit[kMaxExchangeBits];  // global_bits[j] - m
};
// d_ranks: device array of n_ranks slice pointers (2^m amplitudes each).
template <typename T>
void launch_exchange(hipStream_t s, void* const* d_ranks, uint32_t n_ranks, uint32_t m, uint32_t k,
                     const ExchangeBits& bits);
void launch_narrow_f64(hipStream_t s, const void* src_f64, void* dst_f32, uint64_t len);

// ---- fused pass (k_fused.hip) ----
template <typename T>
void launch_fused_pass(hipStream_t s, void* state, uint32_t n, uint32_t m, uint32_t n_stages, uint32_t features,
                       uint32_t reg_bits, const PassDev* pd, const StageDev* stages, const DevOp* ops,
                       const PhaseTerm* terms);

// ---- batched plans (k_batch.hip) ----
struct BatchLaunch {
  uint32_t threads = 512;      // workgroup size: 2^(m_max - G)
  size_t lds_bytes = 0;        // 2^m_max amplitudes
  bool nontemporal = false;
};
template <typename T>
void launch_batch_round(hipStream_t s, const BatchLaunch& cfg, void* pool, const BatchItem* items, uint32_t n_items,
                        const PassDev* passes, const StageDev* stages, const DevOp* ops, const PhaseTerm* terms);
template <typename T>
void launch_batch_persistent(hipStream_t s, const BatchLaunch& cfg, void* pool, const BatchItem* items,
                             uint32_t n_items, uint32_t grid, const PassDev* passes, const StageDev* stages,
                             const DevOp* ops, const PhaseTerm* terms, uint32_t* done, BatchCtl* ctl,
                             uint32_t spin_limit);
// Resident workgroups per CU of the persistent executor.
template <typename T> int batch_persistent_occupancy(const BatchLaunch& cfg);
// Allow the LDS size of `cfg` on both executors (call before graph capture).
template <typename T> void batch_configure(const BatchLaunch& cfg);
// Write every instance's initial state (BatchPrep) and reset the queue.
template <typename T>
void launch_batch_init(hipStream_t s, void* pool, const BatchPrep* prep, uint32_t n_inst, uint32_t* done,
                       BatchCtl* ctl);
void launch_batch_reset(hipStream_t s, uint32_t* done, uint32_t n_inst, BatchCtl* ctl);
// MaxCut expectation per instance: out[i] = sum_x |psi_x|^2 cut_i(x); one
// partial per 4096-amplitude block (descs[i].block_begin numbers them).
constexpr uint32_t kBatchCutBlockBits = 12;
template <typename T>
void launch_batch_cut(hipStream_t s, const void* pool, const CutDesc* descs, uint32_t n_inst, uint32_t n_blocks,
                      double* partials, double* out);

// ---- planner cost model (k_mlp.hip) ----
constexpr int kMlpIn = 16;        // features per candidate pass
constexpr int kMlpHidden = 64;    // three hidden ReLU layers
constexpr int kMlpParams = kMlpIn * kMlpHidden + kMlpHidden + 2 * (kMlpHidden * kMlpHidden + kMlpHidden) +
                           kMlpHidden + 1;
// out[i] = model(X[i, :]) for B rows of kMlpIn floats.
void launch_mlp_forward(hipStream_t s, const float* params, const float* X, uint32_t B, float* out);
// grads = d/dparams of mean((model(X) - y)^2), loss[0] = that mean (both device, overwritten).
void launch_mlp_grad(hipStream_t s, const float* params, const float* X, const float* y, uint32_t B, float* grads,
                     float* loss);
// One Adam step (t counts from 1).
void launch_mlp_adam(hipStream_t s, float* params, const float* grads, float* m, float* v, float lr, float b1,
                     float b2, float eps, uint32_t t);

// ---- readout (k_readout.hip) ----
// Result layout in `out` (device, 3 doubles): [sum of |a|^2 over bit q clear,
// sum over bit q set, non-finite count]; qubit < 0 puts everything in [0].  `partials` needs 3 * kReduceBlocks doubles.
constexpr uint32_t kReduceBlocks = 4096;
template <typename T>
void launch_reduce_branch(hipStream_t s, const void* state, uint32_t n, int32_t qubit, double* partials,
                          double* out);
// out[0] = sum of |a|^4, out[1] = sum of |a|^2, out[2] = non-finite count (same partials as above).
template <typename T>
void launch_sum_p2(hipStream_t s, const void* state, uint32_t n, double* partials, double* out);
// FP64 norm into res (the reduce above), then psi /= sqrt(norm) on the device (no host sync).
template <typename T>
void launch_normalize(hipStream_t s, void* state, uint32_t n, double* partials, double* res);
// Per-block partials (4 doubles each: Re, Im of sum conj(a) b, sum |a-b|^2, max |a-b|^2) of two
// states of equal size; returns the block count (<= kReduceBlocks).
template <typename TA, typename TB>
uint32_t launch_compare(hipStream_t s, const void* a, const void* b, uint32_t n, double* partials);
// Up to kExpectGroup Pauli terms sharing X mask `xmask`, one read of the state:
// out[t] = sum_i (-1)^{|i & zy[t]|} w_t(i), with w_t = |psi_i|^2 when xmask == 0 and
// otherwise Re (use_im[t] == 0) or Im of conj(psi[i ^ xmask]) psi[i] over the
// indices i with xmask's top bit clear.  `partials` needs kExpectGroup * kReduceBlocks
// doubles; `out` kExpectGroup doubles.
constexpr uint32_t kExpectGroup = 16;
template <typename T>
void launch_expect_group(hipStream_t s, const void* state, uint32_t n, uint64_t xmask, const uint64_t* zy,
                         const uint8_t* use_im, uint32_t count, double* partials, double* out);
template <typename T>
void launch_collapse(hipStream_t s, void* state, uint32_t n, uint32_t qubit, int32_t outcome, double scale);

// Marginal histogram over k qubits into `out` (2^k doubles, device, zeroed by the launcher).
template <typename T>
void launch_marginal(hipStream_t s, const void* state, uint32_t n, uint32_t k, const uint32_t* qubits,
                     double* out);
// Sampler: block sums (2^(n - kSampleBlockBits) doubles, scanned in place into
// an exclusive prefix), then one wavefront per shot descends.
constexpr uint32_t kSampleBlockBits = 12;
template <typename T>
void launch_sample(hipStream_t s, const void* state, uint32_t n, double* block_sums, double* scan_tmp,
                   const double* uniforms, uint32_t shots, uint64_t* out);

}  // namespace aleph
