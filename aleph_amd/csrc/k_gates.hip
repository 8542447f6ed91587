// Per-gate kernels: one full-state sweep per gate.  These back the
// Backend::apply_gate entry point, the unfused plan, and dense blocks wider
// than the fused executor's register tile.  CUDA peers: apply_1q,
// apply_kq_tiled, apply_diag, apply_phase_poly, apply_cnot in
// crates/aleph-cuda/src/sv/kernels.cu and diag.cu.
#include "common.hpp"
#include "kernels.hpp"

namespace aleph {
namespace {

constexpr uint32_t kBlock = 256;

// Blocks for one thread per item, capped so that the dispatch stays under
// 2^32 work-items (the HSA packet's limit); kernels whose item count can
// exceed the cap loop grid-stride.
constexpr uint64_t kMaxBlocks = 1ull << 20;
inline uint32_t grid_for(uint64_t items) {
  uint64_t g = (items + kBlock - 1) / kBlock;
  if (g > kMaxBlocks) g = kMaxBlocks;
  return (uint32_t)(g == 0 ? 1 : g);
}

struct M8 { double v[8]; };
struct M32 { double v[32]; };
struct Perm64 { uint32_t p[64]; };

template <typename T>
__global__ void k_init_basis(C2<T>* __restrict__ s, uint64_t len) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += (uint64_t)gridDim.x * blockDim.x) {
    C2<T> v;
    v.x = i == 0 ? (T)1 : (T)0;
    v.y = 0;
    s[i] = v;
  }
}

template <typename T>
__global__ void k_1q(C2<T>* __restrict__ s, uint64_t pairs, uint32_t t, uint64_t cm, M8 m) {
  for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < pairs; p += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t i0 = insert_zero(p, t), i1 = i0 | (1ull << t);
    if ((i0 & cm) != cm) continue;
    const C2<T> x = s[i0], y = s[i1];
    const T a = m.v[0], b = m.v[1], c = m.v[2], d = m.v[3], e = m.v[4], f = m.v[5], g = m.v[6], h = m.v[7];
    C2<T> nx, ny;
    nx.x = a * x.x - b * x.y + c * y.x - d * y.y;
    nx.y = a * x.y + b * x.x + c * y.y + d * y.x;
    ny.x = e * x.x - f * x.y + g * y.x - h * y.y;
    ny.y = e * x.y + f * x.x + g * y.y + h * y.x;
    s[i0] = nx;
    s[i1] = ny;
  }
}

template <typename T>
__global__ void k_2q(C2<T>* __restrict__ s, uint64_t quads, uint32_t qmsb, uint32_t qlsb, uint64_t cm, M32 m) {
  const uint32_t lo = qmsb < qlsb ? qmsb : qlsb, hi = qmsb < qlsb ? qlsb : qmsb;
  for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < quads; p += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t i = insert_zero(insert_zero(p, lo), hi);
    if ((i & cm) != cm) continue;
    const uint64_t bm = 1ull << qmsb, bl = 1ull << qlsb;
    const uint64_t idx[4] = {i, i | bl, i | bm, i | bm | bl};
    C2<T> v[4];
  #pragma unroll
    for (int c = 0; c < 4; c++) v[c] = s[idx[c]];
  #pragma unroll
    for (int r = 0; r < 4; r++) {
      T re = 0, im = 0;
  #pragma unroll
      for (int c = 0; c < 4; c++) {
        const T mr = m.v[2 * (r * 4 + c)], mi = m.v[2 * (r * 4 + c) + 1];
        re += mr * v[c].x - mi * v[c].y;
        im += mr * v[c].y + mi * v[c].x;
      }
      C2<T> o;
      o.x = re;
      o.y = im;
      s[idx[r]] = o;
    }
  }
}

// Dense k-qubit block.  A workgroup stages the matrix in LDS (transposed so
// the lanes of a wavefront read consecutive words) and walks tiles of
// groups; each lane produces one output amplitude of one group.
template <typename T>
__global__ void __launch_bounds__(256) k_kq(C2<T>* __restrict__ s, uint64_t groups, uint32_t k, Perm64 q,
                                            uint64_t cm, const C2<T>* __restrict__ mat) {
  extern __shared__ __align__(16) unsigned char smem[];
  const uint32_t dim = 1u << k;
  C2<T>* mt = reinterpret_cast<C2<T>*>(smem);   // dim*dim, transposed
  C2<T>* buf = mt + dim * dim;                  // 256 staged inputs
  for (uint32_t e = threadIdx.x; e < dim * dim; e += blockDim.x) {
    const uint32_t r = e / dim, c = e % dim;
    mt[c * dim + r] = mat[e];
  }
  // sorted target positions for the group base
  uint32_t sorted[8];
  for (uint32_t j = 0; j < k; j++) sorted[j] = q.p[j];
  for (uint32_t a = 0; a < k; a++)
    for (uint32_t b = a + 1; b < k; b++)
      if (sorted[b] < sorted[a]) { const uint32_t t = sorted[a]; sorted[a] = sorted[b]; sorted[b] = t; }
  const uint32_t per_tile = blockDim.x / dim;   // groups per tile
  const uint32_t lg = threadIdx.x / dim, row = threadIdx.x % dim;
  uint64_t off = 0;                             // offset of matrix index `row`
  for (uint32_t pbit = 0; pbit < k; pbit++)
    if ((row >> pbit) & 1u) off |= 1ull << q.p[k - 1 - pbit];
  __syncthreads();
  for (uint64_t g0 = (uint64_t)blockIdx.x * per_tile; g0 < groups; g0 += (uint64_t)gridDim.x * per_tile) {
    const uint64_t g = g0 + lg;
    uint64_t b = g;
    for (uint32_t j = 0; j < k; j++) b = insert_zero(b, sorted[j]);
    const bool live = g < groups && (b & cm) == cm;
    if (g < groups) buf[threadIdx.x] = s[b | off];
    __syncthreads();
    if (live) {
      T re = 0, im = 0;
      const C2<T>* in = buf + lg * dim;
      for (uint32_t c = 0; c < dim; c++) {
        const C2<T> mv = mt[c * dim + row], v = in[c];
        re += mv.x * v.x - mv.y * v.y;
        im += mv.x * v.y + mv.y * v.x;
      }
      C2<T> o;
      o.x = re;
      o.y = im;
      s[b | off] = o;
    }
    __syncthreads();
  }
}

template <typename T>
__global__ void k_diag(C2<T>* __restrict__ s, uint64_t len, uint32_t k, Perm64 q, uint64_t cm,
                       const double* __restrict__ d) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += (uint64_t)gridDim.x * blockDim.x) {
    if ((i & cm) != cm) continue;
    uint32_t sel = 0;
    for (uint32_t j = 0; j < k; j++) sel = (sel << 1) | (uint32_t)((i >> q.p[j]) & 1ull);
    const T dr = (T)d[2 * sel], di = (T)d[2 * sel + 1];
    const C2<T> x = s[i];
    C2<T> y;
    y.x = x.x * dr - x.y * di;
    y.y = x.x * di + x.y * dr;
    s[i] = y;
  }
}

template <typename T>
__global__ void k_swap(C2<T>* __restrict__ s, uint64_t quads, uint32_t a, uint32_t b, uint64_t cm) {
  const uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
  for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < quads; p += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t i = insert_zero(insert_zero(p, lo), hi);
    if ((i & cm) != cm) continue;
    const uint64_t x = i | (1ull << a), y = i | (1ull << b);
    const C2<T> t = s[x];
    s[x] = s[y];
    s[y] = t;
  }
}

template <typename T>
__global__ void k_phase_poly(C2<T>* __restrict__ s, uint64_t len, uint32_t nt, const double* __restrict__ ang,
                             const uint64_t* __restrict__ conds, const uint32_t* __restrict__ off) {
  for (uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; x < len; x += (uint64_t)gridDim.x * blockDim.x) {
    double phi = 0.0;
    for (uint32_t t = 0; t < nt; t++) {
      bool all = true;
      for (uint32_t c = off[t]; c < off[t + 1]; c++)
        if ((__popcll(conds[c] & x) & 1) == 0) { all = false; break; }
      if (all) phi += ang[t];
    }
    double sn, cs;
    sincos(phi, &sn, &cs);
    const C2<T> v = s[x];
    C2<T> o;
    o.x = v.x * (T)cs - v.y * (T)sn;
    o.y = v.x * (T)sn + v.y * (T)cs;
    s[x] = o;
  }
}

template <typename T>
__global__ void k_permute(const C2<T>* __restrict__ src, C2<T>* __restrict__ dst, uint64_t len, uint32_t n,
                          Perm64 perm) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t j = 0;
    for (uint32_t l = 0; l < n; l++)
      if ((i >> l) & 1ull) j |= 1ull << perm.p[l];
    dst[i] = src[j];
  }
}

__global__ void k_widen(const C2<float>* __restrict__ a, C2<double>* __restrict__ b, uint64_t len) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += (uint64_t)gridDim.x * blockDim.x) {
    C2<double> v;
    v.x = a[i].x;
    v.y = a[i].y;
    b[i] = v;
  }
}

__global__ void k_narrow(const C2<double>* __restrict__ a, C2<float>* __restrict__ b, uint64_t len) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += (uint64_t)gridDim.x * blockDim.x) {
    C2<float> v;
    v.x = (float)a[i].x;
    v.y = (float)a[i].y;
    b[i] = v;
  }
}

Perm64 to_perm(const uint32_t* q, uint32_t k) {
  Perm64 p{};
  for (uint32_t j = 0; j < k && j < 64; j++) p.p[j] = q[j];
  return p;
}

}  // namespace

template <typename T>
void launch_init_basis(hipStream_t st, void* state, uint32_t n) {
  const uint64_t len = 1ull << n;
  hipLaunchKernelGGL(k_init_basis<T>, dim3(grid_for(len)), dim3(kBlock), 0, st, (C2<T>*)state, len);
  ALEPH_HIP_CHECK(hipGetLastError());
}

template <typename T>
void launch_1q(hipStream_t st, void* state, uint32_t n, uint32_t target, uint64_t cmask, const double* m8) {
  M8 m;
  for (int i = 0; i < 8; i++) m.v[i] = m8[i];
  const uint64_t pairs = 1ull << (n - 1);
  hipLaunchKernelGGL(k_1q<T>, dim3(grid_for(pairs)), dim3(kBlock), 0, st, (C2<T>*)state, pairs, target, cmask, m);
  ALEPH_HIP_CHECK(hipGetLastError());
}

template <typename T>
void launch_2q(hipStream_t st, void* state, uint32_t n, uint32_t qm, uint32_t ql, uint64_t cmask, const double* m32) {
  M32 m;
  for (int i = 0; i < 32; i++) m.v[i] = m32[i];
  const uint64_t quads = 1ull << (n - 2);
  hipLaunchKernelGGL(k_2q<T>, dim3(grid_for(quads)), dim3(kBlock), 0, st, (C2<T>*)state, quads, qm, ql, cmask, m);
  ALEPH_HIP_CHECK(hipGetLastError());
}

template <typename T>
void launch_kq(hipStream_t st, void* state, uint32_t n, uint32_t k, const uint32_t* qubits, uint64_t cmask,
               const void* dmat) {
  if (k >= 4 && k <= 6 && n >= k + 4) {
    launch_kq_mfma<T>(st, state, n, k, qubits, cmask, dmat);
    return;
  }
  const uint32_t dim = 1u << k;
  const uint64_t groups = 1ull << (n - k);
  const uint32_t per_tile = kBlock / dim;
  uint64_t blocks = (groups + per_tile - 1) / per_tile;
  if (blocks > 65536) blocks = 65536;
  const size_t lds = sizeof(C2<T>) * ((size_t)dim * dim + kBlock);
  allow_full_lds(reinterpret_cast<const void*>(&k_kq<T>));
  hipLaunchKernelGGL(k_kq<T>, dim3((uint32_t)blocks), dim3(kBlock), lds, st, (C2<T>*)state, groups, k,
                     to_perm(qubits, k), cmask, (const C2<T>*)dmat);
  ALEPH_HIP_CHECK(hipGetLastError());
}

template <typename T>
void launch_diag(hipStream_t st, void* state, uint32_t n, uint32_t k, const uint32_t* qubits, uint64_t cmask,
                 const double* ddiag) {
  const uint64_t len = 1ull << n;
  hipLaunchKernelGGL(k_diag<T>, dim3(grid_for(len)), dim3(kBlock), 0, st, (C2<T>*)state, len, k,
                     to_perm(qubits, k), cmask, ddiag);
  ALEPH_HIP_CHECK(hipGetLastError());
}

template <typename T>
void launch_swap(hipStream_t st, void* state, uint32_t n, uint32_t a, uint32_t b, uint64_t cmask) {
  const uint64_t quads = 1ull << (n - 2);
  hipLaunchKernelGGL(k_swap<T>, dim3(grid_for(quads)), dim3(kBlock), 0, st, (C2<T>*)state, quads, a, b, cmask);
  ALEPH_HIP_CHECK(hipGetLastError());
}

template <typename T>
void launch_phase_poly(hipStream_t st, void* state, uint32_t n, uint32_t nt, const double* angles,
                       const uint64_t* conds, const uint32_t* offsets) {
  const uint64_t len = 1ull << n;
  hipLaunchKernelGGL(k_phase_poly<T>, dim3(grid_for(len)), dim3(kBlock), 0, st, (C2<T>*)state, len, nt, angles,
                     conds, offsets);
  ALEPH_HIP_CHECK(hipGetLastError());
}

template <typename T>
void launch_permute(hipStream_t st, const void* src, void* dst, uint32_t n, const uint32_t* perm) {
  const uint64_t len = 1ull << n;
  hipLaunchKernelGGL(k_permute<T>, dim3(grid_for(len)), dim3(kBlock), 0, st, (const C2<T>*)src, (C2<T>*)dst, len,
                     n, to_perm(perm, n));
  ALEPH_HIP_CHECK(hipGetLastError());
}

void launch_widen_f32(hipStream_t st, const void* src, void* dst, uint64_t len) {
  hipLaunchKernelGGL(k_widen, dim3(grid_for(len)), dim3(kBlock), 0, st, (const C2<float>*)src, (C2<double>*)dst, len);
  ALEPH_HIP_CHECK(hipGetLastError());
}

void launch_narrow_f64(hipStream_t st, const void* src, void* dst, uint64_t len) {
  hipLaunchKernelGGL(k_narrow, dim3(grid_for(len)), dim3(kBlock), 0, st, (const C2<double>*)src, (C2<float>*)dst, len);
  ALEPH_HIP_CHECK(hipGetLastError());
}

#define ALEPH_INST(T)                                                                                   \
  template void launch_init_basis<T>(hipStream_t, void*, uint32_t);                                     \
  template void launch_1q<T>(hipStream_t, void*, uint32_t, uint32_t, uint64_t, const double*);          \
  template void launch_2q<T>(hipStream_t, void*, uint32_t, uint32_t, uint32_t, uint64_t, const double*);\
  template void launch_kq<T>(hipStream_t, void*, uint32_t, uint32_t, const uint32_t*, uint64_t, const void*); \
  template void launch_diag<T>(hipStream_t, void*, uint32_t, uint32_t, const uint32_t*, uint64_t, const double*); \
  template void launch_swap<T>(hipStream_t, void*, uint32_t, uint32_t, uint32_t, uint64_t);             \
  template void launch_phase_poly<T>(hipStream_t, void*, uint32_t, uint32_t, const double*, const uint64_t*, \
                                     const uint32_t*);                                                  \
  template void launch_permute<T>(hipStream_t, const void*, void*, uint32_t, const uint32_t*);
ALEPH_INST(double)
ALEPH_INST(float)
#undef ALEPH_INST

}  // namespace aleph
