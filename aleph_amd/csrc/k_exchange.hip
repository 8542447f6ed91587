// Rank-slice exchange of a distributed state held on one device: global
// physical bit global_bits[j] trades places with local bit m-k+j (the top k
// local bits, so every moved piece is a contiguous chunk of 2^(m-k)
// amplitudes).  Chunk c of rank r and chunk c' of rank r' form an
// involution, so one in-place swap kernel over all pairs does the whole
// exchange: every moved amplitude is read once and written once, with no
// scratch copy (crates/aleph-cuda/src/dist/exchange.rs:133-200 stages each
// pair through a scratch buffer instead).
#include <algorithm>
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
// sigma_R(r) (r <= sigma_R(r)), permutes each through LDS and writes it back
// in the other's place: new[(r, t)] = old[(sigma_R(r), sigma_T(t))].  Every
// amplitude is read once and written once.  LDS slot of tile element t is
// t + t/32: the permuting read gathers lds[sigma_T(t)], and when sigma_T
// sends the low five tile bits to the upper ones, 32 consecutive t would
// otherwise hit LDS addresses 32 elements apart, all in one bank group.
constexpr uint32_t kPermMaxTileBits = 10;
// Both tiles are loaded into registers at once, kPermPerThread elements a
// thread (eight: at n=30 a sweep takes 8.0 ms in FP64 against 8.2 ms at
// four, 12 ms at two and 10 ms at sixteen, where registers limit the waves);
// they then pass through one LDS tile in turn, so a workgroup holds a single
// tile of LDS.
constexpr uint32_t kPermPerThread = 8;
// Every index the kernel needs is a bitwise or of precomputed parts: the
// tile pair's base addresses are ors of per-block-index-bit contributions,
// and element t = threadIdx.x + i * blockDim.x (blockDim a power of two)
// splits into the thread's bits and the wave-uniform bits of i.
struct PermMap {
  uint32_t nt, nb;                      // tile bits; block-index bits
  uint32_t t_pos[kPermMaxTileBits];     // qubit of tile bit j (j < k: qubit j)
  uint32_t t_img[kPermMaxTileBits];     // tile bit that tile bit j moves to
  uint64_t blk_a[40];                   // block-index bit b set: or into tile r's base address
  uint64_t blk_b[40];                   //   ... and into the partner tile sigma_R(r)'s
  uint64_t off_i[kPermPerThread];       // address bits of element i * blockDim.x
  uint32_t src_i[kPermPerThread];       // permuted tile index of element i * blockDim.x
  uint32_t ng;                          // one-qubit gates applied in LDS before permuting
  uint32_t g_bit[kPermMaxGates];        // tile bit each gate acts on
  double g_m[kPermMaxGates][8];
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

__device__ __forceinline__ uint32_t lds_slot(uint32_t t) { return t + (t >> 5); }

// One-qubit gates on tile bits, applied to the tile held in LDS.
template <typename T>
__device__ void tile_gates(T* v, const PermMap& pm, uint32_t tile) {
  for (uint32_t g = 0; g < pm.ng; g++) {
    const uint32_t bit = pm.g_bit[g];
    const double* m = pm.g_m[g];
    for (uint32_t h = threadIdx.x; h < tile / 2; h += blockDim.x) {
      const uint32_t j0 = ((h >> bit) << (bit + 1)) | (h & ((1u << bit) - 1u));
      const uint32_t i0 = lds_slot(j0), i1 = lds_slot(j0 | (1u << bit));
      const double x0 = v[i0].x, y0 = v[i0].y, x1 = v[i1].x, y1 = v[i1].y;
      v[i0].x = m[0] * x0 - m[1] * y0 + m[2] * x1 - m[3] * y1;
      v[i0].y = m[0] * y0 + m[1] * x0 + m[2] * y1 + m[3] * x1;
      v[i1].x = m[4] * x0 - m[5] * y0 + m[6] * x1 - m[7] * y1;
      v[i1].y = m[4] * y0 + m[5] * x0 + m[6] * y1 + m[7] * x1;
    }
    __syncthreads();
  }
}

template <typename T>
__global__ void __launch_bounds__(128) k_bitperm(T* __restrict__ s, PermMap pm) {
  extern __shared__ unsigned char lds_raw[];
  T* lds = reinterpret_cast<T*>(lds_raw);
  const uint32_t tile = 1u << pm.nt;
  uint64_t base_a = 0, base_b = 0;
  for (uint32_t b = 0; b < pm.nb; b++) {
    if ((blockIdx.x >> b) & 1u) {
      base_a |= pm.blk_a[b];
      base_b |= pm.blk_b[b];
    }
  }
  // Addresses order tiles as their numbers do (the tile-number bits sit at
  // increasing qubits), so the pair is handled once, from its lower tile.
  if (base_b < base_a) return;
  const bool pair = base_b != base_a;
  const uint64_t off_t = scatter_bits(threadIdx.x, pm.t_pos, pm.nt);
  const uint32_t src_t = permute_bits(threadIdx.x, pm.t_img, pm.nt);
  // Components in plain arrays: arrays of the vector type stay in scratch
  // memory for FP64 instead of registers.
  using E = decltype(T::x);
  E ax[kPermPerThread], ay[kPermPerThread], bx[kPermPerThread], by[kPermPerThread];
#pragma unroll
  for (uint32_t i = 0; i < kPermPerThread; i++) {
    const uint32_t t = threadIdx.x + i * blockDim.x;
    if (t < tile) {
      const uint64_t off = off_t | pm.off_i[i];
      const T u = s[base_a | off];
      ax[i] = u.x;
      ay[i] = u.y;
      if (pair) {
        const T w = s[base_b | off];
        bx[i] = w.x;
        by[i] = w.y;
      }
    }
  }
  // Tile r through LDS: its permuted values belong to tile r2.
#pragma unroll
  for (uint32_t i = 0; i < kPermPerThread; i++) {
    const uint32_t t = threadIdx.x + i * blockDim.x;
    if (t < tile) {
      lds[lds_slot(t)].x = ax[i];
      lds[lds_slot(t)].y = ay[i];
    }
  }
  __syncthreads();
  tile_gates(lds, pm, tile);
#pragma unroll
  for (uint32_t i = 0; i < kPermPerThread; i++) {
    const uint32_t t = threadIdx.x + i * blockDim.x;
    if (t < tile) {
      const T u = lds[lds_slot(src_t | pm.src_i[i])];
      ax[i] = u.x;
      ay[i] = u.y;
    }
  }
  if (pair) {
    __syncthreads();
#pragma unroll
    for (uint32_t i = 0; i < kPermPerThread; i++) {
      const uint32_t t = threadIdx.x + i * blockDim.x;
      if (t < tile) {
        lds[lds_slot(t)].x = bx[i];
        lds[lds_slot(t)].y = by[i];
      }
    }
    __syncthreads();
    tile_gates(lds, pm, tile);
#pragma unroll
    for (uint32_t i = 0; i < kPermPerThread; i++) {
      const uint32_t t = threadIdx.x + i * blockDim.x;
      if (t < tile) {
        const T u = lds[lds_slot(src_t | pm.src_i[i])];
        bx[i] = u.x;
        by[i] = u.y;
      }
    }
  }
#pragma unroll
  for (uint32_t i = 0; i < kPermPerThread; i++) {
    const uint32_t t = threadIdx.x + i * blockDim.x;
    if (t < tile) {
      const uint64_t off = off_t | pm.off_i[i];
      T u, w;
      u.x = ax[i];
      u.y = ay[i];
      if (pair) {
        w.x = bx[i];
        w.y = by[i];
        s[base_a | off] = w;
        s[base_b | off] = u;
      } else {
        s[base_a | off] = u;
      }
    }
  }
}

}  // namespace

template <typename T>
void launch_bitperm(hipStream_t s, void* state, uint32_t n, const uint32_t* sigma, uint32_t n_gates,
                    const uint32_t* gate_qubits, const double* gate_mats) {
  PermMap pm{};
  const uint32_t k = std::min<uint32_t>(5, n);
  uint64_t tmask = 0;
  for (uint32_t q = 0; q < k; q++) tmask |= (1ull << q) | (1ull << sigma[q]);
  // Tile bits: the low k qubits first (consecutive amplitudes), then their images.
  for (uint32_t q = 0; q < k; q++) pm.t_pos[pm.nt++] = q;
  for (uint32_t q = k; q < n; q++)
    if ((tmask >> q) & 1ull) pm.t_pos[pm.nt++] = q;
  uint32_t r_pos[64], nr = 0;
  for (uint32_t q = 0; q < n; q++)
    if (!((tmask >> q) & 1ull)) r_pos[nr++] = q;
  auto index_of = [](const uint32_t* pos, uint32_t count, uint32_t q) {
    for (uint32_t j = 0; j < count; j++)
      if (pos[j] == q) return j;
    return count;
  };
  for (uint32_t j = 0; j < pm.nt; j++) pm.t_img[j] = index_of(pm.t_pos, pm.nt, sigma[pm.t_pos[j]]);
  // Workgroups that run together take consecutive block indices, so the low
  // block-index bits go to the tile-number bits whose higher qubit, over the
  // bit and its image, is lowest: both tiles of neighbouring workgroups then
  // differ only in low address bits (the same pages), for the QFT's reversal
  // the middle qubits first.  With the tile number taken straight from the
  // block index, the partner tiles of neighbouring workgroups of the reversal
  // lie 2^(n-6) amplitudes apart.
  uint32_t order[64];
  for (uint32_t j = 0; j < nr; j++) order[j] = j;
  std::stable_sort(order, order + nr, [&](uint32_t x, uint32_t y) {
    return std::max(r_pos[x], sigma[r_pos[x]]) < std::max(r_pos[y], sigma[r_pos[y]]);
  });
  if (nr > 40) throw Failure{ALEPH_HIP_ERR_INVALID_ARGUMENT, "bitperm: state too wide"};
  pm.nb = nr;
  for (uint32_t b = 0; b < nr; b++) {
    const uint32_t q = r_pos[order[b]];
    pm.blk_a[b] = 1ull << q;
    pm.blk_b[b] = 1ull << sigma[q];
  }
  using V = typename std::conditional<std::is_same<T, double>::value, double2, float2>::type;
  const size_t tile = (size_t)1 << pm.nt;
  const uint32_t threads = std::max<uint32_t>(64, (uint32_t)(tile / kPermPerThread));
  for (uint32_t i = 0; i < kPermPerThread; i++) {
    const uint32_t t = i * threads;
    for (uint32_t j = 0; j < pm.nt; j++) {
      if ((t >> j) & 1u) {
        pm.off_i[i] |= 1ull << pm.t_pos[j];
        pm.src_i[i] |= 1u << pm.t_img[j];
      }
    }
  }
  if (n_gates > kPermMaxGates) throw Failure{ALEPH_HIP_ERR_INVALID_ARGUMENT, "bitperm: too many gates"};
  pm.ng = n_gates;
  for (uint32_t g = 0; g < n_gates; g++) {
    pm.g_bit[g] = index_of(pm.t_pos, pm.nt, gate_qubits[g]);
    if (pm.g_bit[g] >= pm.nt) throw Failure{ALEPH_HIP_ERR_INVALID_ARGUMENT, "bitperm: gate off the tile"};
    for (int e = 0; e < 8; e++) pm.g_m[g][e] = gate_mats[8 * g + e];
  }
  const size_t lds = (tile + tile / 32 + 1) * sizeof(V);
  hipLaunchKernelGGL(k_bitperm<V>, dim3(1u << pm.nb), dim3(threads), lds, s, (V*)state, pm);
  ALEPH_HIP_CHECK(hipGetLastError());
}

template void launch_bitperm<double>(hipStream_t, void*, uint32_t, const uint32_t*, uint32_t, const uint32_t*,
                                    const double*);
template void launch_bitperm<float>(hipStream_t, void*, uint32_t, const uint32_t*, uint32_t, const uint32_t*,
                                   const double*);

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
