// Rank-slice exchange of a distributed state held on one device: global
// physical bit global_bits[j] trades places with local bit m-k+j (the top k
// local bits, so every moved piece is a contiguous chunk of 2^(m-k)
// amplitudes).  Chunk c of rank r and chunk c' of rank r' form an
// involution, so one in-place swap kernel over all pairs does the whole
// exchange: every moved amplitude is read once and written once, with no
// scratch copy (crates/aleph-cuda/src/dist/exchange.rs:133-200 stages each
// pair through a scratch buffer instead).
#include "common.hpp"
#include "kernels.hpp"

namespace aleph {
namespace {

template <typename T>
__global__ void __launch_bounds__(256) k_exchange(T* const* ranks, uint32_t m, uint32_t k, ExchangeBits bits,
                                                  uint64_t total) {
  const uint64_t chunk_bits = m - k;
  const uint64_t chunk_mask = (1ull << chunk_bits) - 1ull;
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t r = (uint32_t)(e >> m);
    const uint64_t local = e & ((1ull << m) - 1ull);
    const uint32_t c = (uint32_t)(local >> chunk_bits);
    uint32_t r2 = r, c2 = 0;
    for (uint32_t j = 0; j < k; j++) {
      const uint32_t rb = bits.rank_bit[j];
      c2 |= ((r >> rb) & 1u) << j;
      r2 = (r2 & ~(1u << rb)) | (((c >> j) & 1u) << rb);
    }
    // Each unordered pair once, from its smaller (rank, chunk) side.
    if (r2 < r || (r2 == r && c2 <= c)) continue;
    const uint64_t off = local & chunk_mask;
    T* a = ranks[r] + ((uint64_t)c << chunk_bits) + off;
    T* b = ranks[r2] + ((uint64_t)c2 << chunk_bits) + off;
    const T va = *a, vb = *b;
    *a = vb;
    *b = va;
  }
}

}  // namespace

template <typename T>
void launch_exchange(hipStream_t s, void* const* d_ranks, uint32_t n_ranks, uint32_t m, uint32_t k,
                     const ExchangeBits& bits) {
  const uint64_t total = (uint64_t)n_ranks << m;
  const uint64_t blocks = std::min<uint64_t>((total + 255) / 256, 1u << 20);
  using V = typename std::conditional<std::is_same<T, double>::value, double2, float2>::type;
  hipLaunchKernelGGL(k_exchange<V>, dim3((uint32_t)std::max<uint64_t>(blocks, 1)), dim3(256), 0, s,
                     (V* const*)d_ranks, m, k, bits, total);
  ALEPH_HIP_CHECK(hipGetLastError());
}

template void launch_exchange<double>(hipStream_t, void* const*, uint32_t, uint32_t, uint32_t, const ExchangeBits&);
template void launch_exchange<float>(hipStream_t, void* const*, uint32_t, uint32_t, uint32_t, const ExchangeBits&);

}  // namespace aleph
