// Dense k-qubit blocks (k = 4, 5, 6) on the matrix cores of gfx950.
//
// A block U (D x D complex, D = 2^k, row-major, qubits[0] = MSB of the
// matrix index) acts on every group of D amplitudes that differ only in the
// k target bits: Y = U X with X the D x groups matrix of the state.  In real
// form Y' = R X' with R = [[Re U, -Im U], [Im U, Re U]] (2D x 2D) and X'
// the real parts stacked over the imaginary parts, which is a real GEMM of
// M = K = 2D against N = groups; each wavefront computes N = 16 groups per
// tile with v_mfma_{f64,f32}_16x16x4 (2D/16 row tiles x 2D/4 K steps).
// The reference applies these blocks with a scalar gather/matvec/scatter
// per group (crates/aleph-sv/src/kernels/unitary_kq.rs:127-175) and on CUDA
// with per-thread rows (crates/aleph-cuda/src/sv/kernels.cu:226-283).
//
// Operand layouts of the 16x16x4 instructions (lane l of 64, value v of 4),
// as measured on gfx950:
//   A (16 x 4): row l % 16, column l / 16;   B (4 x 16): row l / 16, column l % 16;
//   C (16 x 16): column l % 16, row 4 (l / 16) + v for f32, row (l / 16) + 4 v for f64.
// So lane l feeds group (l % 16) of the tile, K index 4 s + l / 16 at step s,
// and owns outputs of that group at rows 16 r + (those rows).
//
// The R fragments are built once per workgroup in LDS in fragment order
// (one conflict-free read per lane per MFMA), together with the index offset
// of every matrix index t.
#include <utility>

#include "common.hpp"
#include "kernels.hpp"

namespace aleph {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
  using Acc = d4;
  static __device__ __forceinline__ Acc fma(double a, double b, Acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
};
template <>
struct Mfma<float> {
  using Acc = f4;
  static __device__ __forceinline__ Acc fma(float a, float b, Acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
};

struct KqArgs {
  uint32_t msb_first[8];   // target qubits, qubits[0] = MSB of the matrix index
  uint32_t sorted[8];      // the same, ascending (group base construction)
};

template <typename T, int K>
__global__ void __launch_bounds__(256) k_kq_mfma(C2<T>* __restrict__ s, uint64_t groups, KqArgs q, uint64_t cm,
                                                 const C2<T>* __restrict__ mat) {
  constexpr int D = 1 << K;
  constexpr int MT = 2 * D / 16;   // row tiles of R
  constexpr int KS = 2 * D / 4;    // K steps
  constexpr int NA = D / 4;        // input amplitudes per lane
  using M = Mfma<T>;
  using Acc = typename M::Acc;
  extern __shared__ __align__(16) unsigned char smem[];
  T* frag = reinterpret_cast<T*>(smem);                                   // MT * KS * 64
  uint64_t* dep = reinterpret_cast<uint64_t*>(frag + (size_t)MT * KS * 64);  // D offsets

  for (uint32_t e = threadIdx.x; e < (uint32_t)(MT * KS * 64); e += blockDim.x) {
    const uint32_t l = e & 63, rs = e >> 6;
    const uint32_t r = rs / KS, st = rs % KS;
    const uint32_t row = 16 * r + (l & 15), col = 4 * st + (l >> 4);
    const C2<T> u = mat[(size_t)(row % D) * D + (col % D)];
    T v;
    if (row < (uint32_t)D) v = col < (uint32_t)D ? u.x : -u.y;
    else v = col < (uint32_t)D ? u.y : u.x;
    frag[e] = v;
  }
  for (uint32_t t = threadIdx.x; t < (uint32_t)D; t += blockDim.x) {
    uint64_t o = 0;
#pragma unroll
    for (int j = 0; j < K; j++)
      if ((t >> j) & 1u) o |= 1ull << q.msb_first[K - 1 - j];
    dep[t] = o;
  }
  __syncthreads();

  const uint32_t lane = threadIdx.x & 63, kq = lane >> 4, col = lane & 15;
  const uint64_t tiles = (groups + 15) / 16;
  const uint64_t waves = (uint64_t)gridDim.x * (blockDim.x >> 6);
  for (uint64_t tile = (uint64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); tile < tiles; tile += waves) {
    const uint64_t g = tile * 16 + col;
    const bool in = g < groups;
    uint64_t base = in ? g : 0;
#pragma unroll
    for (int j = 0; j < K; j++) base = insert_zero(base, q.sorted[j]);
    const bool live = in && (base & cm) == cm;
    C2<T> x[NA];
#pragma unroll
    for (int a = 0; a < NA; a++) {
      if (in) x[a] = s[base | dep[4 * a + kq]];
      else x[a].x = x[a].y = 0;
    }
    Acc acc[MT];
#pragma unroll
    for (int r = 0; r < MT; r++) acc[r] = Acc{0, 0, 0, 0};
#pragma unroll
    for (int st = 0; st < KS; st++) {
      const T b = st < NA ? x[st].x : x[st - NA].y;
#pragma unroll
      for (int r = 0; r < MT; r++) acc[r] = M::fma(frag[((size_t)r * KS + st) * 64 + lane], b, acc[r]);
    }
    if (live) {
#pragma unroll
      for (int r = 0; r < MT / 2; r++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
          C2<T> o;
          o.x = acc[r][v];
          o.y = acc[r + MT / 2][v];
          const uint32_t row = sizeof(T) == 8 ? kq + 4 * v : 4 * kq + v;
          s[base | dep[16 * r + row]] = o;
        }
    }
  }
}

template <typename T, int K>
void launch_k(hipStream_t st, void* state, uint32_t n, const uint32_t* qubits, uint64_t cmask, const void* dmat) {
  constexpr int D = 1 << K;
  const size_t lds = sizeof(T) * (size_t)(2 * D / 16) * (2 * D / 4) * 64 + sizeof(uint64_t) * D;
  allow_full_lds(reinterpret_cast<const void*>(&k_kq_mfma<T, K>));
  KqArgs a{};
  for (int j = 0; j < K; j++) a.msb_first[j] = a.sorted[j] = qubits[j];
  for (int i = 0; i < K; i++)
    for (int j = i + 1; j < K; j++)
      if (a.sorted[j] < a.sorted[i]) std::swap(a.sorted[i], a.sorted[j]);
  const uint64_t groups = 1ull << (n - K);
  const uint64_t tiles = (groups + 15) / 16;
  uint64_t blocks = (tiles + 3) / 4;
  if (blocks > 2048) blocks = 2048;
  hipLaunchKernelGGL((k_kq_mfma<T, K>), dim3((uint32_t)blocks), dim3(256), lds, st, (C2<T>*)state, groups, a, cmask,
                     (const C2<T>*)dmat);
  ALEPH_HIP_CHECK(hipGetLastError());
}

}  // namespace

template <typename T>
void launch_kq_mfma(hipStream_t st, void* state, uint32_t n, uint32_t k, const uint32_t* qubits, uint64_t cmask,
                    const void* dmat) {
  switch (k) {
    case 4: launch_k<T, 4>(st, state, n, qubits, cmask, dmat); break;
    case 5: launch_k<T, 5>(st, state, n, qubits, cmask, dmat); break;
    case 6: launch_k<T, 6>(st, state, n, qubits, cmask, dmat); break;
    default: throw Failure{ALEPH_HIP_ERR_INVALID_ARGUMENT, "matrix-core blocks need k in [4, 6]"};
  }
}

template void launch_kq_mfma<double>(hipStream_t, void*, uint32_t, uint32_t, const uint32_t*, uint64_t, const void*);
template void launch_kq_mfma<float>(hipStream_t, void*, uint32_t, uint32_t, const uint32_t*, uint64_t, const void*);

}  // namespace aleph
