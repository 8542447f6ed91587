// Noise trajectories: one workgroup runs one whole trajectory at a time,
// gate by gate and channel by channel, then draws its terminal outcome.
// States of up to 13 qubits live in LDS (one read of the program from L2,
// no HBM traffic per gate); wider states live in the workgroup's own slice
// of a device buffer.  The RNG is counter-based, so every thread of the
// workgroup computes the same draw without communication.
#include "common.hpp"
#include "noise.hpp"

namespace aleph {
namespace {

constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ull;

// splitmix64 finaliser (Steele, Lea & Flood, OOPSLA 2014).
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Per-shot key: crates/aleph-sv/src/noise/apply.rs:169-180.
__device__ __forceinline__ uint64_t shot_key(uint64_t seed, uint64_t shot) {
  return mix64(seed * kGolden + shot + kGolden);
}

// Draw c (0-based) of the splitmix64 stream seeded with `key`, as a
// 53-bit uniform in [0, 1).
__device__ __forceinline__ double uniform(uint64_t key, uint32_t c) {
  return (double)(mix64(key + (uint64_t)(c + 1) * kGolden) >> 11) * (1.0 / 9007199254740992.0);
}

__device__ __forceinline__ uint64_t insert_zero(uint64_t x, uint32_t b) {
  return ((x >> b) << (b + 1)) | (x & ((1ull << b) - 1ull));
}

// z * i^e for e in 0..3 (exact: component swaps and negations).
__device__ __forceinline__ double2 mul_ipow(double2 z, uint32_t e) {
  switch (e & 3u) {
    case 0: return z;
    case 1: return make_double2(-z.y, z.x);
    case 2: return make_double2(-z.x, -z.y);
    default: return make_double2(z.y, -z.x);
  }
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// Sum over the workgroup; every thread returns the same value (the wave
// partials are added in one fixed order by every thread).
__device__ double block_sum(double v, double* red) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double t = 0.0;
  const uint32_t nw = (blockDim.x + 63u) >> 6;
  for (uint32_t w = 0; w < nw; w++) t += red[w];
  __syncthreads();
  return t;
}

template <int K>
__device__ void apply_dense(double2* s, uint32_t n, const TrajOp& op, const double* m) {
  constexpr uint32_t D = 1u << K;
  uint32_t sorted[K];
#pragma unroll
  for (int i = 0; i < K; i++) sorted[i] = op.q[i];
#pragma unroll
  for (int i = 0; i < K; i++)
#pragma unroll
    for (int j = 0; j + 1 < K - i; j++)
      if (sorted[j] > sorted[j + 1]) {
        const uint32_t t = sorted[j];
        sorted[j] = sorted[j + 1];
        sorted[j + 1] = t;
      }
  const uint64_t groups = 1ull << (n - K);
  for (uint64_t g = threadIdx.x; g < groups; g += blockDim.x) {
    uint64_t base = g;
#pragma unroll
    for (int i = 0; i < K; i++) base = insert_zero(base, sorted[i]);
    if ((base & op.cmask) != op.cmask) continue;
    double2 a[D];
    uint64_t idx[D];
#pragma unroll
    for (uint32_t j = 0; j < D; j++) {
      uint64_t x = base;
#pragma unroll
      for (int t = 0; t < K; t++)
        if ((j >> (K - 1 - t)) & 1u) x |= 1ull << op.q[t];
      idx[j] = x;
      a[j] = s[x];
    }
#pragma unroll
    for (uint32_t r = 0; r < D; r++) {
      double re = 0.0, im = 0.0;
#pragma unroll
      for (uint32_t c = 0; c < D; c++) {
        const double mr = m[2 * (r * D + c)], mi = m[2 * (r * D + c) + 1];
        re += mr * a[c].x - mi * a[c].y;
        im += mr * a[c].y + mi * a[c].x;
      }
      s[idx[r]] = make_double2(re, im);
    }
  }
}

__device__ void apply_diag(double2* s, uint32_t n, const TrajOp& op, const double* d) {
  const uint64_t N = 1ull << n;
  for (uint64_t x = threadIdx.x; x < N; x += blockDim.x) {
    if ((x & op.cmask) != op.cmask) continue;
    uint32_t j = 0;
    for (uint32_t t = 0; t < op.k; t++) j = (j << 1) | (uint32_t)((x >> op.q[t]) & 1ull);
    s[x] = cmul(s[x], make_double2(d[2 * j], d[2 * j + 1]));
  }
}

// DiagonalPhase (crates/aleph-ir/src/diagonal_phase.rs:60-72): term t adds
// angles[t] when popcount(cond & x) is odd for each of its conditions.
__device__ void apply_phase(double2* s, uint32_t n, const TrajOp& op, const double* angles, const uint64_t* conds,
                            const uint32_t* offsets) {
  const uint64_t N = 1ull << n;
  for (uint64_t x = threadIdx.x; x < N; x += blockDim.x) {
    double phi = 0.0;
    for (uint32_t t = 0; t < op.count; t++) {
      bool fire = true;
      for (uint32_t c = offsets[t]; c < offsets[t + 1]; c++)
        if ((__popcll(conds[c] & x) & 1) == 0) fire = false;
      if (fire) phi += angles[t];
    }
    double sn, cs;
    sincos(phi, &sn, &cs);
    s[x] = cmul(s[x], make_double2(cs, sn));
  }
}

// Pauli string i^ny X^xm Z^zm (Y = i X Z on its qubit):
// psi'[x ^ xm] = i^ny (-1)^popcount(zm & x) psi[x].
__device__ void apply_pauli(double2* s, uint32_t n, uint64_t xm, uint64_t zm) {
  const uint32_t ny = (uint32_t)__popcll(xm & zm);
  const uint64_t N = 1ull << n;
  if (xm == 0) {
    for (uint64_t x = threadIdx.x; x < N; x += blockDim.x)
      s[x] = mul_ipow(s[x], ny + 2u * ((uint32_t)__popcll(zm & x) & 1u));
    return;
  }
  const uint32_t lb = (uint32_t)__ffsll((long long)xm) - 1u;
  for (uint64_t t = threadIdx.x; t < N / 2; t += blockDim.x) {
    const uint64_t x = insert_zero(t, lb), y = x ^ xm;
    const double2 a = s[x], b = s[y];
    s[y] = mul_ipow(a, ny + 2u * ((uint32_t)__popcll(zm & x) & 1u));
    s[x] = mul_ipow(b, ny + 2u * ((uint32_t)__popcll(zm & y) & 1u));
  }
}

// One-qubit quantum jump (crates/aleph-sv/src/noise/apply.rs:68-133).
// Branch probabilities are summed only until the drawn branch is known.
__device__ void apply_kraus(double2* s, uint32_t n, const TrajOp& op, const double* kraus, double u, double* red) {
  const uint32_t q = op.q[0];
  const uint64_t qb = 1ull << q, half = (1ull << n) / 2;
  double acc = 0.0, pc = 0.0;
  uint32_t chosen = op.count;
  for (uint32_t b = 0; b < op.count; b++) {
    const double* k = kraus + 8 * b;
    double part = 0.0;
    for (uint64_t t = threadIdx.x; t < half; t += blockDim.x) {
      const uint64_t i0 = insert_zero(t, q), i1 = i0 | qb;
      const double2 a0 = s[i0], a1 = s[i1];
      const double o0r = k[0] * a0.x - k[1] * a0.y + k[2] * a1.x - k[3] * a1.y;
      const double o0i = k[0] * a0.y + k[1] * a0.x + k[2] * a1.y + k[3] * a1.x;
      const double o1r = k[4] * a0.x - k[5] * a0.y + k[6] * a1.x - k[7] * a1.y;
      const double o1i = k[4] * a0.y + k[5] * a0.x + k[6] * a1.y + k[7] * a1.x;
      part += o0r * o0r + o0i * o0i + o1r * o1r + o1i * o1i;
    }
    const double pb = block_sum(part, red);
    acc += pb;
    pc = pb;
    if (u < acc) {
      chosen = b;
      break;
    }
  }
  if (chosen == op.count) chosen = op.count - 1;   // the last branch absorbs rounding (pc is its probability)
  if (pc < 1e-300) return;                          // degenerate branch: keep the pre-channel state
  const double inv = 1.0 / sqrt(pc);
  const double* k = kraus + 8 * chosen;
  for (uint64_t t = threadIdx.x; t < half; t += blockDim.x) {
    const uint64_t i0 = insert_zero(t, q), i1 = i0 | qb;
    const double2 a0 = s[i0], a1 = s[i1];
    const double o0r = k[0] * a0.x - k[1] * a0.y + k[2] * a1.x - k[3] * a1.y;
    const double o0i = k[0] * a0.y + k[1] * a0.x + k[2] * a1.y + k[3] * a1.x;
    const double o1r = k[4] * a0.x - k[5] * a0.y + k[6] * a1.x - k[7] * a1.y;
    const double o1i = k[4] * a0.y + k[5] * a0.x + k[6] * a1.y + k[7] * a1.x;
    s[i0] = make_double2(o0r * inv, o0i * inv);
    s[i1] = make_double2(o1r * inv, o1i * inv);
  }
}

template <bool kLds>
__global__ void __launch_bounds__(kLds ? kNoiseThreadsLds : kNoiseThreadsWide) k_trajectories(TrajProgram p, double2* slices, uint64_t* out) {
  extern __shared__ double2 lds_state[];
  __shared__ double red[16];
  __shared__ double scan[16];
  __shared__ unsigned long long found, last_pos;
  const uint32_t n = p.n;
  const uint64_t N = 1ull << n;
  double2* s = kLds ? lds_state : slices + (size_t)blockIdx.x * N;
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6, nw = (blockDim.x + 63u) >> 6;
  for (uint64_t shot = blockIdx.x; shot < p.shots; shot += gridDim.x) {
    const uint64_t key = shot_key(p.seed, p.shot0 + shot);
    uint32_t draw = 0;
    for (uint64_t i = threadIdx.x; i < N; i += blockDim.x) s[i] = make_double2(i == 0 ? 1.0 : 0.0, 0.0);
    if (threadIdx.x == 0) {
      found = ~0ull;
      last_pos = 0;
    }
    __syncthreads();
    for (uint32_t o = 0; o < p.n_ops; o++) {
      const TrajOp& op = p.ops[o];
      switch (op.kind) {
        case TrajOp::DENSE:
          if (op.k == 1) apply_dense<1>(s, n, op, p.reals + op.off);
          else if (op.k == 2) apply_dense<2>(s, n, op, p.reals + op.off);
          else apply_dense<3>(s, n, op, p.reals + op.off);
          break;
        case TrajOp::DIAG:
          apply_diag(s, n, op, p.reals + op.off);
          break;
        case TrajOp::PHASE:
          apply_phase(s, n, op, p.reals + op.off, p.masks + op.off2, p.index + op.off3);
          break;
        case TrajOp::PAULI: {
          // crates/aleph-sv/src/noise/apply.rs:46-66: one draw, cumulative
          // weights, the last term absorbs rounding.
          const double u = uniform(key, draw++);
          const double* w = p.reals + op.off;
          uint32_t chosen = op.count - 1;
          double acc = 0.0;
          for (uint32_t b = 0; b < op.count; b++) {
            acc += w[b];
            if (u < acc) {
              chosen = b;
              break;
            }
          }
          const uint64_t xm = p.masks[op.off2 + 2 * chosen], zm = p.masks[op.off2 + 2 * chosen + 1];
          if (xm | zm) apply_pauli(s, n, xm, zm);
          break;
        }
        default:
          apply_kraus(s, n, op, p.reals + op.off, uniform(key, draw++), red);
          break;
      }
      __syncthreads();
    }
    // Terminal draw (crates/aleph-sv/src/sampling.rs:61-91): the first index
    // whose running sum of |psi|^2 exceeds u * total, zero-mass bins
    // skipped, the last positive bin when rounding leaves nothing.  Each
    // thread owns a contiguous segment; a wave scan gives its offset.
    const uint64_t seg = N >= blockDim.x ? N / blockDim.x : 1ull;
    const uint64_t lo = (uint64_t)threadIdx.x * seg, hi = lo < N ? lo + seg : lo;
    double part = 0.0;
    uint64_t mine_last = 0;
    bool any = false;
    for (uint64_t i = lo; i < hi; i++) {
      const double2 a = s[i];
      const double pr = a.x * a.x + a.y * a.y;
      if (pr > 0.0) {
        part += pr;
        mine_last = i;
        any = true;
      }
    }
    double incl = part;
    for (int d = 1; d < 64; d <<= 1) {
      const double y = __shfl_up(incl, d);
      if ((int)lane >= d) incl += y;
    }
    double excl = __shfl_up(incl, 1);
    if (lane == 0) excl = 0.0;
    if (lane == 63) scan[wave] = incl;
    if (any) atomicMax(&last_pos, (unsigned long long)mine_last);
    __syncthreads();
    double woff = 0.0, total = 0.0;
    for (uint32_t w = 0; w < nw; w++) {
      if (w < wave) woff += scan[w];
      total += scan[w];
    }
    const double target = uniform(key, draw++) * total;
    double cum = woff + excl;
    if (cum <= target) {
      for (uint64_t i = lo; i < hi; i++) {
        const double2 a = s[i];
        const double pr = a.x * a.x + a.y * a.y;
        if (pr > 0.0) {
          cum += pr;
          if (target < cum) {
            atomicMin(&found, (unsigned long long)i);
            break;
          }
        }
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      uint64_t idx = found != ~0ull ? (uint64_t)found : (uint64_t)last_pos;
      // Readout error (apply.rs:135-158): one draw per qubit that has one,
      // in ascending qubit order, flipping the bit with P(1-t | t).
      for (uint32_t q = 0; q < n; q++) {
        if (!((p.readout_mask >> q) & 1ull)) continue;
        const uint32_t bit = (uint32_t)((idx >> q) & 1ull);
        if (uniform(key, draw++) < p.readout[2 * q + bit]) idx ^= 1ull << q;
      }
      out[shot] = idx;
    }
    __syncthreads();
  }
}

}  // namespace

void trajectory_geometry(uint32_t n, uint32_t shots, size_t budget, uint32_t* workgroups, size_t* bytes) {
  if (n <= kNoiseLdsMaxQubits) {
    *workgroups = std::max<uint32_t>(1u, std::min<uint32_t>(shots, 16384u));
    *bytes = 0;
    return;
  }
  const size_t slice = (size_t{1} << n) * sizeof(double2);
  const size_t fit = std::max<size_t>(1, budget / slice);
  *workgroups = (uint32_t)std::max<size_t>(1, std::min<size_t>({(size_t)shots, fit, 1024}));
  *bytes = (size_t)*workgroups * slice;
}

void launch_trajectories(hipStream_t s, const TrajProgram& p, uint32_t workgroups, void* slices, uint64_t* out) {
  if (p.shots == 0) return;
  if (p.n <= kNoiseLdsMaxQubits) {
    const uint32_t threads = std::min<uint32_t>(kNoiseThreadsLds, std::max<uint32_t>(64u, (1u << p.n) / 2u));
    const size_t lds = ((size_t)1 << p.n) * sizeof(double2);
    // The largest state (13 qubits, 128 KiB) plus the kernel's static LDS fits the CU's 160 KiB.
    allow_full_lds((const void*)k_trajectories<true>, (int)((size_t{1} << kNoiseLdsMaxQubits) * sizeof(double2)));
    hipLaunchKernelGGL(k_trajectories<true>, dim3(workgroups), dim3(threads), lds, s, p, (double2*)nullptr, out);
  } else {
    hipLaunchKernelGGL(k_trajectories<false>, dim3(workgroups), dim3(kNoiseThreadsWide), 0, s, p, (double2*)slices,
                       out);
  }
  ALEPH_HIP_CHECK(hipGetLastError());
}

}  // namespace aleph
