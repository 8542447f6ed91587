// Rank-slice exchange of a distributed state held on one device: global
// physical bit global_bits[j] trades places with local bit m-k+j (the top k
// local bits, so every moved piece is a contiguous chunk of 2^(m-k)
// amplitudes).  Chunk c of rank r and chunk c' of rank r' form an
// involution, so one in-place swap kernel over all pairs does the whole
// exchange: every moved amplitude is read once and written once, with no
// scratch copy (crates/aleph-cuda/src/dist/exchange.rs:133-200 stages each
// pair through a scratch buffer instead).
#include <type_traits>

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

// An involutive qubit permutation sigma, applied in place.  Tile bits T are
// the low k qubits and their images (closed under sigma), so a tile holds
// runs of 2^k consecutive amplitudes; the other qubits R index tiles, and
// sigma maps tile r to tile sigma_R(r).  One workgroup loads tiles r and
// sigma_R(r) (r <= sigma_R(r)) into LDS and writes each back permuted:
// new[(r, t)] = old[(sigma_R(r), sigma_T(t))].  Every amplitude is read once
// and written once.
constexpr uint32_t kPermMaxTileBits = 10;
struct PermMap {
  uint32_t nt, nr;
  uint32_t t_pos[kPermMaxTileBits];     // qubit of tile bit j (j < k: qubit j)
  uint32_t t_img[kPermMaxTileBits];     // tile bit that tile bit j moves to
  uint32_t r_pos[40];
  uint32_t r_img[40];
};

__device__ __forceinline__ uint64_t scatter_bits(uint64_t v, const uint32_t* pos, uint32_t count) {
  uint64_t x = 0;
  for (uint32_t j = 0; j < count; j++) x |= ((v >> j) & 1ull) << pos[j];
  return x;
}

__device__ __forceinline__ uint32_t permute_bits(uint32_t v, const uint32_t* img, uint32_t count) {
  uint32_t x = 0;
  for (uint32_t j = 0; j < count; j++) x |= ((v >> j) & 1u) << img[j];
  return x;
}

template <typename T>
__global__ void __launch_bounds__(256) k_bitperm(T* __restrict__ s, PermMap pm) {
  extern __shared__ unsigned char lds_raw[];
  T* a = reinterpret_cast<T*>(lds_raw);
  const uint32_t tile = 1u << pm.nt;
  T* b = a + tile;
  const uint32_t r = blockIdx.x;
  uint32_t r2 = 0;
  for (uint32_t j = 0; j < pm.nr; j++) r2 |= ((r >> j) & 1u) << pm.r_img[j];
  if (r2 < r) return;
  const uint64_t base_a = scatter_bits(r, pm.r_pos, pm.nr), base_b = scatter_bits(r2, pm.r_pos, pm.nr);
  for (uint32_t t = threadIdx.x; t < tile; t += blockDim.x) {
    const uint64_t off = scatter_bits(t, pm.t_pos, pm.nt);
    a[t] = s[base_a | off];
    if (r2 != r) b[t] = s[base_b | off];
  }
  __syncthreads();
  for (uint32_t t = threadIdx.x; t < tile; t += blockDim.x) {
    const uint64_t off = scatter_bits(t, pm.t_pos, pm.nt);
    const uint32_t src = permute_bits(t, pm.t_img, pm.nt);
    if (r2 == r) {
      s[base_a | off] = a[src];
    } else {
      s[base_a | off] = b[src];
      s[base_b | off] = a[src];
    }
  }
}

}  // namespace

template <typename T>
void launch_bitperm(hipStream_t s, void* state, uint32_t n, const uint32_t* sigma) {
  PermMap pm{};
  const uint32_t k = std::min<uint32_t>(5, n);
  uint64_t tmask = 0;
  for (uint32_t q = 0; q < k; q++) tmask |= (1ull << q) | (1ull << sigma[q]);
  // Tile bits: the low k qubits first (consecutive amplitudes), then their images.
  for (uint32_t q = 0; q < k; q++) pm.t_pos[pm.nt++] = q;
  for (uint32_t q = k; q < n; q++)
    if ((tmask >> q) & 1ull) pm.t_pos[pm.nt++] = q;
  for (uint32_t q = 0; q < n; q++)
    if (!((tmask >> q) & 1ull)) pm.r_pos[pm.nr++] = q;
  auto index_of = [](const uint32_t* pos, uint32_t count, uint32_t q) {
    for (uint32_t j = 0; j < count; j++)
      if (pos[j] == q) return j;
    return count;
  };
  for (uint32_t j = 0; j < pm.nt; j++) pm.t_img[j] = index_of(pm.t_pos, pm.nt, sigma[pm.t_pos[j]]);
  for (uint32_t j = 0; j < pm.nr; j++) pm.r_img[j] = index_of(pm.r_pos, pm.nr, sigma[pm.r_pos[j]]);
  using V = typename std::conditional<std::is_same<T, double>::value, double2, float2>::type;
  const size_t lds = 2 * ((size_t)1 << pm.nt) * sizeof(V);
  hipLaunchKernelGGL(k_bitperm<V>, dim3(1u << pm.nr), dim3(256), lds, s, (V*)state, pm);
  ALEPH_HIP_CHECK(hipGetLastError());
}

template void launch_bitperm<double>(hipStream_t, void*, uint32_t, const uint32_t*);
template void launch_bitperm<float>(hipStream_t, void*, uint32_t, const uint32_t*);

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
