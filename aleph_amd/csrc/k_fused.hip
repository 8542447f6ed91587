// Fused LDS-staged pass executor for gfx950.
//
// One workgroup = one chunk of 2^m amplitudes (the index bits outside the
// pass's chunk qubits are fixed by blockIdx).  The chunk is read from HBM
// once, every op of the pass is applied while it sits in registers (2^G
// amplitudes per thread) or is re-distributed through LDS between stages,
// and it is written back once: one HBM round trip for the whole pass,
// instead of one per gate as in the reference's per-gate launches
// (crates/aleph-cuda/src/sv/backend.rs:515-581).
#include "common.hpp"
#include "fused_chunk.hpp"
#include "kernels.hpp"

namespace aleph {
namespace {

using namespace fused;

template <typename T, int G, int BT, int FEAT, bool NT>
__global__ void __launch_bounds__(BT, 2) k_fused_pass(C2<T>* __restrict__ state, const PassDev* __restrict__ pd,
                                                      const StageDev* __restrict__ stages,
                                                      const DevOp* __restrict__ ops,
                                                      const PhaseTerm* __restrict__ terms) {
  extern __shared__ __align__(16) unsigned char smem[];
  fused_chunk<T, G, FEAT, NT, false>(state, pd, stages, ops, terms, blockIdx.x, reinterpret_cast<C2<T>*>(smem));
}

}  // namespace

template <typename T, int G, int BT, int FEAT, bool NT>
static void launch_bt(hipStream_t stream, void* state, uint32_t n, uint32_t m, uint32_t n_stages, const PassDev* pd,
                      const StageDev* stages, const DevOp* ops, const PhaseTerm* terms) {
  const uint64_t blocks = 1ull << (n - m);
  const uint32_t threads = 1u << (m - G);
  const size_t lds = n_stages > 1 ? (sizeof(C2<T>) << m) : 0;
  allow_full_lds(reinterpret_cast<const void*>(&k_fused_pass<T, G, BT, FEAT, NT>));
  hipLaunchKernelGGL((k_fused_pass<T, G, BT, FEAT, NT>), dim3((uint32_t)blocks), dim3(threads), lds, stream,
                     reinterpret_cast<C2<T>*>(state), pd, stages, ops, terms);
  ALEPH_HIP_CHECK(hipGetLastError());
}

template <typename T, int G, int FEAT, bool NT>
static void launch_threads(hipStream_t stream, void* state, uint32_t n, uint32_t m, uint32_t n_stages,
                           const PassDev* pd, const StageDev* stages, const DevOp* ops, const PhaseTerm* terms) {
  const uint32_t threads = 1u << (m - G);
  if (threads <= 64) launch_bt<T, G, 64, FEAT, NT>(stream, state, n, m, n_stages, pd, stages, ops, terms);
  else if (threads <= 256) launch_bt<T, G, 256, FEAT, NT>(stream, state, n, m, n_stages, pd, stages, ops, terms);
  else launch_bt<T, G, 512, FEAT, NT>(stream, state, n, m, n_stages, pd, stages, ops, terms);
}

template <typename T, int G, bool NT>
static void launch_feat(hipStream_t stream, void* state, uint32_t n, uint32_t m, uint32_t n_stages, uint32_t feat,
                        const PassDev* pd, const StageDev* stages, const DevOp* ops, const PhaseTerm* terms) {
  switch (feat & (kFeat2Q | kFeatPhase)) {
    case 0: launch_threads<T, G, 0, NT>(stream, state, n, m, n_stages, pd, stages, ops, terms); break;
    case kFeat2Q: launch_threads<T, G, kFeat2Q, NT>(stream, state, n, m, n_stages, pd, stages, ops, terms); break;
    case kFeatPhase:
      launch_threads<T, G, kFeatPhase, NT>(stream, state, n, m, n_stages, pd, stages, ops, terms);
      break;
    default:
      launch_threads<T, G, kFeat2Q | kFeatPhase, NT>(stream, state, n, m, n_stages, pd, stages, ops, terms);
      break;
  }
}

template <typename T>
void launch_fused_pass(hipStream_t stream, void* state, uint32_t n, uint32_t m, uint32_t n_stages, uint32_t feat,
                       uint32_t reg_bits, const PassDev* pd, const StageDev* stages, const DevOp* ops,
                       const PhaseTerm* terms) {
  const bool nt = (feat & kFeatNontemporal) != 0;
  if (reg_bits == 5) {
    if (nt) launch_feat<T, 5, true>(stream, state, n, m, n_stages, feat, pd, stages, ops, terms);
    else launch_feat<T, 5, false>(stream, state, n, m, n_stages, feat, pd, stages, ops, terms);
  } else {
    if (nt) launch_feat<T, 4, true>(stream, state, n, m, n_stages, feat, pd, stages, ops, terms);
    else launch_feat<T, 4, false>(stream, state, n, m, n_stages, feat, pd, stages, ops, terms);
  }
}

template void launch_fused_pass<double>(hipStream_t, void*, uint32_t, uint32_t, uint32_t, uint32_t, uint32_t,
                                        const PassDev*, const StageDev*, const DevOp*, const PhaseTerm*);
template void launch_fused_pass<float>(hipStream_t, void*, uint32_t, uint32_t, uint32_t, uint32_t, uint32_t,
                                       const PassDev*, const StageDev*, const DevOp*, const PhaseTerm*);

}  // namespace aleph
