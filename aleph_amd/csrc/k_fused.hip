// Fused LDS-staged pass executor for gfx950.
//
// One workgroup = one chunk of 2^m amplitudes (the index bits outside the
// pass's chunk qubits are fixed by blockIdx).  The chunk is read from HBM
// once, every op of the pass is applied while it sits in registers (16
// amplitudes per thread) or is re-distributed through LDS between stages,
// and it is written back once: one HBM round trip for the whole pass,
// instead of one per gate as in the reference's per-gate launches
// (crates/aleph-cuda/src/sv/backend.rs:515-581).
#include "common.hpp"
#include "kernels.hpp"
#include "plan.hpp"

namespace aleph {
namespace {

// XOR-fold swizzle of the chunk-local index into LDS slots: bits >= SW are
// folded into the low SW bits, so lanes whose thread bits map to high local
// bits still spread over distinct 128-byte bank windows.
template <int SW>
__device__ __forceinline__ uint32_t swz(uint32_t li) {
  const uint32_t h = li >> SW;
  return li ^ ((h ^ (h >> SW) ^ (h >> (2 * SW)) ^ (h >> (3 * SW))) & ((1u << SW) - 1u));
}

template <typename T>
__device__ __forceinline__ C2<T> cmac2(T ar, T ai, C2<T> x, T br, T bi, C2<T> y) {
  C2<T> o;
  o.x = ar * x.x - ai * x.y + br * y.x - bi * y.y;
  o.y = ar * x.y + ai * x.x + br * y.y + bi * y.x;
  return o;
}

template <typename T, int B>
__device__ __forceinline__ void op_1q(C2<T> (&a)[kRegAmps], const DevOp* __restrict__ op, uint64_t tgo,
                                      const uint64_t (&roff)[kRegAmps]) {
  const T m00r = (T)op->m[0], m00i = (T)op->m[1], m01r = (T)op->m[2], m01i = (T)op->m[3];
  const T m10r = (T)op->m[4], m10i = (T)op->m[5], m11r = (T)op->m[6], m11i = (T)op->m[7];
  const uint64_t cm = op->cmask;
  if (cm == 0) {
#pragma unroll
    for (int r = 0; r < kRegAmps; r++) {
      if (r & (1 << B)) continue;
      const int s = r | (1 << B);
      const C2<T> x = a[r], y = a[s];
      a[r] = cmac2(m00r, m00i, x, m01r, m01i, y);
      a[s] = cmac2(m10r, m10i, x, m11r, m11i, y);
    }
  } else {
#pragma unroll
    for (int r = 0; r < kRegAmps; r++) {
      if (r & (1 << B)) continue;
      const int s = r | (1 << B);
      const C2<T> x = a[r], y = a[s];
      const bool on = ((tgo | roff[r]) & cm) == cm;
      const C2<T> nx = cmac2(m00r, m00i, x, m01r, m01i, y);
      const C2<T> ny = cmac2(m10r, m10i, x, m11r, m11i, y);
      a[r] = on ? nx : x;
      a[s] = on ? ny : y;
    }
  }
}

template <typename T, int A, int B>  // A = matrix MSB register bit, B = LSB, A > B
__device__ __forceinline__ void op_2q(C2<T> (&a)[kRegAmps], const DevOp* __restrict__ op, uint64_t tgo,
                                      const uint64_t (&roff)[kRegAmps]) {
  T mr[16], mi[16];
#pragma unroll
  for (int e = 0; e < 16; e++) {
    mr[e] = (T)op->m[2 * e];
    mi[e] = (T)op->m[2 * e + 1];
  }
  const uint64_t cm = op->cmask;
#pragma unroll
  for (int r = 0; r < kRegAmps; r++) {
    if (r & ((1 << A) | (1 << B))) continue;
    const int idx[4] = {r, r | (1 << B), r | (1 << A), r | (1 << A) | (1 << B)};
    C2<T> v[4], o[4];
#pragma unroll
    for (int c = 0; c < 4; c++) v[c] = a[idx[c]];
#pragma unroll
    for (int row = 0; row < 4; row++) {
      T re = 0, im = 0;
#pragma unroll
      for (int c = 0; c < 4; c++) {
        re += mr[row * 4 + c] * v[c].x - mi[row * 4 + c] * v[c].y;
        im += mr[row * 4 + c] * v[c].y + mi[row * 4 + c] * v[c].x;
      }
      o[row].x = re;
      o[row].y = im;
    }
    if (cm == 0) {
#pragma unroll
      for (int c = 0; c < 4; c++) a[idx[c]] = o[c];
    } else {
      const bool on = ((tgo | roff[r]) & cm) == cm;
#pragma unroll
      for (int c = 0; c < 4; c++) a[idx[c]] = on ? o[c] : v[c];
    }
  }
}

template <typename T>
__device__ __forceinline__ void op_diag(C2<T> (&a)[kRegAmps], const DevOp* __restrict__ op, uint64_t tgo,
                                        const uint64_t (&roff)[kRegAmps]) {
  const uint64_t ma = op->ma, mb = op->mb, cm = op->cmask;
  const T d0r = (T)op->m[0], d0i = (T)op->m[1], d1r = (T)op->m[2], d1i = (T)op->m[3];
  const T d2r = (T)op->m[4], d2i = (T)op->m[5], d3r = (T)op->m[6], d3i = (T)op->m[7];
#pragma unroll
  for (int r = 0; r < kRegAmps; r++) {
    const uint64_t idx = tgo | roff[r];
    const bool hi = (idx & ma) != 0, lo = (idx & mb) != 0;
    const T dr = hi ? (lo ? d3r : d2r) : (lo ? d1r : d0r);
    const T di = hi ? (lo ? d3i : d2i) : (lo ? d1i : d0i);
    const C2<T> x = a[r];
    C2<T> y;
    y.x = x.x * dr - x.y * di;
    y.y = x.x * di + x.y * dr;
    a[r] = ((idx & cm) == cm) ? y : x;
  }
}

template <typename T>
__device__ __forceinline__ void op_sign(C2<T> (&a)[kRegAmps], const DevOp* __restrict__ op, uint64_t tgo) {
  const uint64_t mo = op->cmask;
  const uint32_t rm = op->a;
  const bool th = (tgo & mo) == mo;
#pragma unroll
  for (int r = 0; r < kRegAmps; r++) {
    const bool flip = th && ((uint32_t)r & rm) == rm;
    a[r].x = flip ? -a[r].x : a[r].x;
    a[r].y = flip ? -a[r].y : a[r].y;
  }
}

template <typename T>
__device__ __forceinline__ void op_phase(C2<T> (&a)[kRegAmps], const DevOp* __restrict__ op,
                                         const PhaseTerm* __restrict__ terms, uint64_t tgo) {
  // Terms touching at most two register bits fold into 11 accumulators
  // (constant, four linear, six pair coefficients); the rare terms on three
  // or more register bits (class 11) are stored after them and are
  // evaluated per amplitude.
  double c0 = 0.0;
  double l0 = 0.0, l1 = 0.0, l2 = 0.0, l3 = 0.0;
  double p01 = 0.0, p02 = 0.0, p03 = 0.0, p12 = 0.0, p13 = 0.0, p23 = 0.0;
  const uint32_t end = op->term_off + op->n_terms;
  const uint32_t gen_end = end + op->b;
  for (uint32_t t = op->term_off; t < end; t++) {
    const PhaseTerm* pt = terms + t;
    const uint64_t mo = pt->mask_out;
    const double add = ((tgo & mo) == mo) ? pt->angle : 0.0;
    switch (pt->cls) {
      case 0: c0 += add; break;
      case 1: l0 += add; break;
      case 2: l1 += add; break;
      case 3: l2 += add; break;
      case 4: l3 += add; break;
      case 5: p01 += add; break;
      case 6: p02 += add; break;
      case 7: p03 += add; break;
      case 8: p12 += add; break;
      case 9: p13 += add; break;
      default: p23 += add; break;
    }
  }
#pragma unroll
  for (int r = 0; r < kRegAmps; r++) {
    double phi = c0;
    if (r & 1) phi += l0;
    if (r & 2) phi += l1;
    if (r & 4) phi += l2;
    if (r & 8) phi += l3;
    if ((r & 3) == 3) phi += p01;
    if ((r & 5) == 5) phi += p02;
    if ((r & 9) == 9) phi += p03;
    if ((r & 6) == 6) phi += p12;
    if ((r & 10) == 10) phi += p13;
    if ((r & 12) == 12) phi += p23;
    for (uint32_t t = end; t < gen_end; t++) {
      const PhaseTerm* pt = terms + t;
      const uint64_t mo = pt->mask_out;
      const uint32_t rm = pt->rmask;
      if ((tgo & mo) == mo && ((uint32_t)r & rm) == rm) phi += pt->angle;
    }
    double s, c;
    sincos(phi, &s, &c);
    const C2<T> x = a[r];
    a[r].x = x.x * (T)c - x.y * (T)s;
    a[r].y = x.x * (T)s + x.y * (T)c;
  }
}

template <typename T, int FEAT>
__device__ __forceinline__ void run_ops(C2<T> (&a)[kRegAmps], const DevOp* __restrict__ ops,
                                        const PhaseTerm* __restrict__ terms, uint32_t begin, uint32_t end,
                                        uint64_t tgo, const uint64_t (&roff)[kRegAmps]) {
  for (uint32_t o = begin; o < end; o++) {
    const DevOp* op = ops + o;
    const uint32_t type = op->type;
    if constexpr ((FEAT & kFeat2Q) == 0) {
      if (type == OP_2Q) continue;
    }
    if constexpr ((FEAT & kFeatPhase) == 0) {
      if (type == OP_PHASE) continue;
    }
    switch (type) {
      case OP_1Q:
        switch (op->a) {
          case 0: op_1q<T, 0>(a, op, tgo, roff); break;
          case 1: op_1q<T, 1>(a, op, tgo, roff); break;
          case 2: op_1q<T, 2>(a, op, tgo, roff); break;
          default: op_1q<T, 3>(a, op, tgo, roff); break;
        }
        break;
      case OP_2Q:
        if constexpr ((FEAT & kFeat2Q) != 0) switch (op->a * 4 + op->b) {
          case 1 * 4 + 0: op_2q<T, 1, 0>(a, op, tgo, roff); break;
          case 2 * 4 + 0: op_2q<T, 2, 0>(a, op, tgo, roff); break;
          case 3 * 4 + 0: op_2q<T, 3, 0>(a, op, tgo, roff); break;
          case 2 * 4 + 1: op_2q<T, 2, 1>(a, op, tgo, roff); break;
          case 3 * 4 + 1: op_2q<T, 3, 1>(a, op, tgo, roff); break;
          default: op_2q<T, 3, 2>(a, op, tgo, roff); break;
        }
        break;
      case OP_DIAG: op_diag<T>(a, op, tgo, roff); break;
      case OP_SIGN: op_sign<T>(a, op, tgo); break;
      default:
        if constexpr ((FEAT & kFeatPhase) != 0) op_phase<T>(a, op, terms, tgo);
        break;
    }
  }
}

template <typename T, int SW, int BT, int FEAT>
__global__ void __launch_bounds__(BT, BT == 256 ? 2 : 1) k_fused_pass(C2<T>* __restrict__ state, const PassDev* __restrict__ pd,
                                                     const StageDev* __restrict__ stages,
                                                     const DevOp* __restrict__ ops,
                                                     const PhaseTerm* __restrict__ terms) {
  extern __shared__ __align__(16) unsigned char smem[];
  C2<T>* lds = reinterpret_cast<C2<T>*>(smem);
  const uint32_t m = pd->m;
  uint64_t base = blockIdx.x;
  for (uint32_t i = 0; i < m; i++) base = insert_zero(base, pd->local_pos[i]);
  const uint32_t tid = threadIdx.x;
  const uint32_t ns = pd->n_stages;
  C2<T> a[kRegAmps];
  for (uint32_t s = 0; s < ns; s++) {
    const StageDev* sd = stages + pd->stage_begin + s;
    uint64_t tgo = base;
    uint32_t tlo = 0;
    const uint32_t ntb = sd->n_thread_bits;
    for (uint32_t j = 0; j < ntb; j++) {
      if ((tid >> j) & 1u) {
        tgo |= sd->tg[j];
        tlo |= sd->tl[j];
      }
    }
    const uint64_t rg0 = sd->rg[0], rg1 = sd->rg[1], rg2 = sd->rg[2], rg3 = sd->rg[3];
    const uint32_t rl0 = sd->rl[0], rl1 = sd->rl[1], rl2 = sd->rl[2], rl3 = sd->rl[3];
    uint64_t roff[kRegAmps];
    uint32_t loff[kRegAmps];
#pragma unroll
    for (int r = 0; r < kRegAmps; r++) {
      roff[r] = ((r & 1) ? rg0 : 0) | ((r & 2) ? rg1 : 0) | ((r & 4) ? rg2 : 0) | ((r & 8) ? rg3 : 0);
      loff[r] = ((r & 1) ? rl0 : 0) | ((r & 2) ? rl1 : 0) | ((r & 4) ? rl2 : 0) | ((r & 8) ? rl3 : 0);
    }
    if (s == 0) {
#pragma unroll
      for (int r = 0; r < kRegAmps; r++) a[r] = state[tgo | roff[r]];
    } else {
#pragma unroll
      for (int r = 0; r < kRegAmps; r++) a[r] = lds[swz<SW>(tlo | loff[r])];
    }
    run_ops<T, FEAT>(a, ops, terms, sd->op_begin, sd->op_end, tgo, roff);
    if (s + 1 == ns) {
#pragma unroll
      for (int r = 0; r < kRegAmps; r++) state[tgo | roff[r]] = a[r];
    } else {
      if (s > 0) __syncthreads();
#pragma unroll
      for (int r = 0; r < kRegAmps; r++) lds[swz<SW>(tlo | loff[r])] = a[r];
      __syncthreads();
    }
  }
}

}  // namespace

template <typename T, int BT, int FEAT>
static void launch_bt(hipStream_t stream, void* state, uint32_t n, uint32_t m, uint32_t n_stages, const PassDev* pd,
                      const StageDev* stages, const DevOp* ops, const PhaseTerm* terms) {
  constexpr int SW = sizeof(T) == 8 ? 3 : 4;
  const uint64_t blocks = 1ull << (n - m);
  const uint32_t threads = 1u << (m - kRegBits);
  const size_t lds = n_stages > 1 ? (sizeof(C2<T>) << m) : 0;
  static bool configured = false;
  if (!configured) {
    ALEPH_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_fused_pass<T, SW, BT, FEAT>),
                                        hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
    configured = true;
  }
  hipLaunchKernelGGL((k_fused_pass<T, SW, BT, FEAT>), dim3((uint32_t)blocks), dim3(threads), lds, stream,
                     reinterpret_cast<C2<T>*>(state), pd, stages, ops, terms);
  ALEPH_HIP_CHECK(hipGetLastError());
}

template <typename T, int FEAT>
static void launch_feat(hipStream_t stream, void* state, uint32_t n, uint32_t m, uint32_t n_stages, const PassDev* pd,
                        const StageDev* stages, const DevOp* ops, const PhaseTerm* terms) {
  const uint32_t threads = 1u << (m - kRegBits);
  if (threads <= 256) launch_bt<T, 256, FEAT>(stream, state, n, m, n_stages, pd, stages, ops, terms);
  else launch_bt<T, 512, FEAT>(stream, state, n, m, n_stages, pd, stages, ops, terms);
}

template <typename T>
void launch_fused_pass(hipStream_t stream, void* state, uint32_t n, uint32_t m, uint32_t n_stages, uint32_t feat,
                       const PassDev* pd, const StageDev* stages, const DevOp* ops, const PhaseTerm* terms) {
  switch (feat & (kFeat2Q | kFeatPhase)) {
    case 0: launch_feat<T, 0>(stream, state, n, m, n_stages, pd, stages, ops, terms); break;
    case kFeat2Q: launch_feat<T, kFeat2Q>(stream, state, n, m, n_stages, pd, stages, ops, terms); break;
    case kFeatPhase: launch_feat<T, kFeatPhase>(stream, state, n, m, n_stages, pd, stages, ops, terms); break;
    default: launch_feat<T, kFeat2Q | kFeatPhase>(stream, state, n, m, n_stages, pd, stages, ops, terms); break;
  }
}

template void launch_fused_pass<double>(hipStream_t, void*, uint32_t, uint32_t, uint32_t, uint32_t, const PassDev*,
                                        const StageDev*, const DevOp*, const PhaseTerm*);
template void launch_fused_pass<float>(hipStream_t, void*, uint32_t, uint32_t, uint32_t, uint32_t, const PassDev*,
                                       const StageDev*, const DevOp*, const PhaseTerm*);

}  // namespace aleph
