// Readout reductions: norm / branch probability, collapse, Pauli
// expectation, marginals and shot sampling.  All accumulation is FP64
// (also for complex64 states, as crates/aleph-cuda/src/sv/readout_f32.cu:1-8
// requires); per-thread sums are Neumaier-compensated and combined with
// error-free two-sums across lanes and workgroups.
#include "common.hpp"
#include "kernels.hpp"

namespace aleph {
namespace {

constexpr uint32_t kThreads = 256;

struct KSum {
  double s, c;
};

__device__ __forceinline__ void kadd(KSum& k, double x) {
  const double t = k.s + x;
  if (fabs(k.s) >= fabs(x)) k.c += (k.s - t) + x;
  else k.c += (x - t) + k.s;
  k.s = t;
}

__device__ __forceinline__ KSum kmerge(KSum a, KSum b) {
  KSum r = a;
  kadd(r, b.s);
  r.c += b.c;
  return r;
}

__device__ __forceinline__ KSum wave_reduce(KSum v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    KSum o;
    o.s = __shfl_down(v.s, off, 64);
    o.c = __shfl_down(v.c, off, 64);
    v = kmerge(v, o);
  }
  return v;
}

// Block-wide reduction of up to 3 compensated sums; result valid in thread 0.
template <int N>
__device__ void block_reduce(KSum (&v)[N]) {
  __shared__ KSum sh[N][kThreads / 64];
  const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i < N; i++) v[i] = wave_reduce(v[i]);
  if (lane == 0)
#pragma unroll
    for (int i = 0; i < N; i++) sh[i][wid] = v[i];
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < N; i++) {
      KSum acc = sh[i][0];
      for (uint32_t w = 1; w < blockDim.x / 64; w++) acc = kmerge(acc, sh[i][w]);
      v[i] = acc;
    }
  }
}

// Grid-stride sweeps take kUnroll amplitudes per thread per iteration, all
// loads issued before any arithmetic, so each wave keeps several 16-byte
// loads in flight.
constexpr int kUnroll = 4;

// MODE 0: [sum over bit q clear, sum over bit q set, non-finite count] of p = |a|^2.
// MODE 1: [sum of p^2, sum of p, non-finite count] (collision probability).
template <typename T, int MODE>
__global__ void __launch_bounds__(kThreads) k_reduce_branch(const C2<T>* __restrict__ s, uint64_t len,
                                                            uint64_t qmask, double* __restrict__ partials) {
  KSum v[3] = {{0, 0}, {0, 0}, {0, 0}};
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i0 < len; i0 += kUnroll * stride) {
    C2<T> a[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; u++) {
      const uint64_t i = i0 + u * stride;
      if (i < len) a[u] = s[i];
      else a[u].x = a[u].y = 0;
    }
#pragma unroll
    for (int u = 0; u < kUnroll; u++) {
      const uint64_t i = i0 + u * stride;
      const double p = (double)a[u].x * (double)a[u].x + (double)a[u].y * (double)a[u].y;
      if (!isfinite(p)) {
        v[2].s += 1.0;
        continue;
      }
      if (MODE == 0) {
        if (i & qmask) kadd(v[1], p);
        else kadd(v[0], p);
      } else {
        kadd(v[0], p * p);
        kadd(v[1], p);
      }
    }
  }
  block_reduce<3>(v);
  if (threadIdx.x == 0) {
    partials[3 * blockIdx.x + 0] = v[0].s + v[0].c;
    partials[3 * blockIdx.x + 1] = v[1].s + v[1].c;
    partials[3 * blockIdx.x + 2] = v[2].s;
  }
}

// Sums `count` interleaved groups of `width` partials into out[0..width).
__global__ void __launch_bounds__(kThreads) k_reduce_final(const double* __restrict__ partials, uint32_t count,
                                                           uint32_t width, double* __restrict__ out) {
  KSum v[3] = {{0, 0}, {0, 0}, {0, 0}};
  for (uint32_t b = threadIdx.x; b < count; b += blockDim.x)
    for (uint32_t w = 0; w < width && w < 3; w++) kadd(v[w], partials[width * b + w]);
  block_reduce<3>(v);
  if (threadIdx.x == 0)
    for (uint32_t w = 0; w < width && w < 3; w++) out[w] = v[w].s + v[w].c;
}

template <typename T>
__global__ void k_collapse(C2<T>* __restrict__ s, uint64_t len, uint64_t qmask, int32_t outcome, double scale) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += (uint64_t)gridDim.x * blockDim.x) {
    const bool keep = ((i & qmask) != 0) == (outcome != 0);
    C2<T> a = s[i];
    if (keep) {
      a.x = (T)((double)a.x * scale);
      a.y = (T)((double)a.y * scale);
    } else {
      a.x = 0;
      a.y = 0;
    }
    s[i] = a;
  }
}

// All Pauli terms that share one X mask `x` in one read of the state.
// With z = conj(psi[i ^ x]) psi[i] and s_t(i) = (-1)^{|i & zy_t|}, term t
// needs R_t = sum_i s_t(i) Re z(i) and I_t = sum_i s_t(i) Im z(i).  For
// x != 0 the sum runs over pairs {i, j = i ^ x} (i with x's top bit clear),
// each amplitude read once: z(j) = conj z(i) and s_t(j) = c_t s_t(i) with
// c_t = (-1)^{|x & zy_t|}, so the pair adds 2 s_t(i) Re z(i) to R_t when
// c_t = 1 (I_t gets 0) and 2 s_t(i) Im z(i) to I_t when c_t = -1 (R_t
// gets 0).  Hence one accumulator per term: acc_t = sum s_t(i) w_t(i) with
// w_t = Re z or Im z; the factor 2 and the placement are applied on the host.
// For x = 0, w = |psi_i|^2 over every i.
constexpr int kGroupTerms = (int)kExpectGroup;
struct GroupTerms {
  uint64_t zy[kGroupTerms];
  uint32_t use_im;   // bit t: term t accumulates Im z
  uint32_t count;
};

template <typename T>
__global__ void __launch_bounds__(kThreads) k_expect_group(const C2<T>* __restrict__ s, uint64_t pairs, uint64_t x,
                                                           uint32_t hbit, GroupTerms g,
                                                           double* __restrict__ partials) {
  double acc[kGroupTerms];
#pragma unroll
  for (int t = 0; t < kGroupTerms; t++) acc[t] = 0.0;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t lo_mask = (1ull << hbit) - 1ull;
  for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < pairs; k += stride) {
    // k enumerates indices with bit hbit clear (every index when x == 0).
    const uint64_t i = x ? (((k & ~lo_mask) << 1) | (k & lo_mask)) : k;
    const C2<T> a = s[i];
    double re, im;
    if (x) {
      const C2<T> b = s[i ^ x];
      re = (double)b.x * (double)a.x + (double)b.y * (double)a.y;
      im = (double)b.x * (double)a.y - (double)b.y * (double)a.x;
    } else {
      re = (double)a.x * (double)a.x + (double)a.y * (double)a.y;
      im = 0.0;
    }
#pragma unroll
    for (int t = 0; t < kGroupTerms; t++) {
      if (t < (int)g.count) {
        const double w = ((g.use_im >> t) & 1u) ? im : re;
        acc[t] += (__popcll(i & g.zy[t]) & 1) ? -w : w;
      }
    }
  }
  // Block reduction per term: wave shuffles, then the block's waves in LDS.
  __shared__ double sh[kGroupTerms][kThreads / 64];
  const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6;
#pragma unroll
  for (int t = 0; t < kGroupTerms; t++) {
    double v = acc[t];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if (lane == 0) sh[t][wid] = v;
  }
  __syncthreads();
  if (threadIdx.x < (uint32_t)g.count) {
    double v = 0.0;
    for (uint32_t w = 0; w < blockDim.x / 64; w++) v += sh[threadIdx.x][w];
    partials[(uint64_t)blockIdx.x * kGroupTerms + threadIdx.x] = v;
  }
}

// Sums `count` rows of `width` partials column-wise into out[0..width)
// (compensated), one workgroup.
__global__ void __launch_bounds__(kThreads) k_reduce_columns(const double* __restrict__ partials, uint32_t count,
                                                             uint32_t width, double* __restrict__ out) {
  __shared__ double sh[kThreads];
  for (uint32_t w = 0; w < width; w++) {
    KSum v = {0, 0};
    for (uint32_t b = threadIdx.x; b < count; b += blockDim.x) kadd(v, partials[(uint64_t)width * b + w]);
    v = wave_reduce(v);
    if ((threadIdx.x & 63u) == 0) sh[threadIdx.x >> 6] = v.s + v.c;
    __syncthreads();
    if (threadIdx.x == 0) {
      KSum acc = {0, 0};
      for (uint32_t i = 0; i < blockDim.x / 64; i++) kadd(acc, sh[i]);
      out[w] = acc.s + acc.c;
    }
    __syncthreads();
  }
}

struct Qubits64 { uint32_t q[64]; };

// LDS-privatised histogram (k <= 12); wider marginals go straight to global atomics.
template <typename T>
__global__ void __launch_bounds__(kThreads) k_marginal(const C2<T>* __restrict__ s, uint64_t len, uint32_t k,
                                                       Qubits64 qs, double* __restrict__ out) {
  extern __shared__ __align__(16) unsigned char smem[];
  double* h = reinterpret_cast<double*>(smem);
  const uint32_t bins = 1u << k;
  const bool priv = k <= 12;
  if (priv) {
    for (uint32_t b = threadIdx.x; b < bins; b += blockDim.x) h[b] = 0.0;
    __syncthreads();
  }
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += (uint64_t)gridDim.x * blockDim.x) {
    const C2<T> a = s[i];
    const double p = (double)a.x * (double)a.x + (double)a.y * (double)a.y;
    uint32_t bin = 0;
    for (uint32_t j = 0; j < k; j++) bin |= (uint32_t)((i >> qs.q[j]) & 1ull) << j;
    if (priv) atomicAdd(&h[bin], p);
    else atomicAdd(&out[bin], p);
  }
  if (priv) {
    __syncthreads();
    for (uint32_t b = threadIdx.x; b < bins; b += blockDim.x)
      if (h[b] != 0.0) atomicAdd(&out[b], h[b]);
  }
}

__global__ void k_zero(double* __restrict__ p, uint64_t count) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = 0.0;
}

// ---- sampler ----
constexpr uint32_t kSampleBlock = 1u << kSampleBlockBits;

template <typename T>
__global__ void __launch_bounds__(kThreads) k_block_sums(const C2<T>* __restrict__ s, uint64_t len,
                                                         double* __restrict__ sums, double* __restrict__ prefix) {
  const uint64_t base = (uint64_t)blockIdx.x * kSampleBlock;
  const uint32_t per = kSampleBlock / kThreads;
  KSum v[1] = {{0, 0}};
  const uint64_t lim = len < base + kSampleBlock ? len : base + kSampleBlock;
  for (uint32_t j = 0; j < per; j++) {
    const uint64_t i = base + (uint64_t)j * kThreads + threadIdx.x;
    if (i < lim) {
      const C2<T> a = s[i];
      kadd(v[0], (double)a.x * (double)a.x + (double)a.y * (double)a.y);
    }
  }
  block_reduce<1>(v);
  if (threadIdx.x == 0) {
    const double t = v[0].s + v[0].c;
    sums[blockIdx.x] = t;
    prefix[blockIdx.x] = t;
  }
}

// Exclusive prefix of `count` sums in place, one workgroup of 1024 threads;
// the grand total lands in total[0].
__global__ void __launch_bounds__(1024) k_scan(double* __restrict__ v, uint64_t count, double* __restrict__ total) {
  __shared__ double part[1024];
  const uint32_t t = threadIdx.x, nt = blockDim.x;
  const uint64_t per = (count + nt - 1) / nt;
  const uint64_t b = t * per, e = b + per < count ? b + per : count;
  double acc = 0.0;
  for (uint64_t i = b; i < e; i++) acc += v[i];
  part[t] = acc;
  __syncthreads();
  if (t == 0) {
    double run = 0.0;
    for (uint32_t j = 0; j < nt; j++) {
      const double x = part[j];
      part[j] = run;
      run += x;
    }
    total[0] = run;
  }
  __syncthreads();
  double run = part[t];
  for (uint64_t i = b; i < e; i++) {
    const double x = v[i];
    v[i] = run;
    run += x;
  }
}

// One wavefront per shot: binary search over block prefixes, then a
// lane-parallel scan inside the chosen block.  Zero-probability amplitudes
// are never returned (strict `target < cumulative`, as
// crates/aleph-sv/src/sampling.rs:61-83 does).
template <typename T>
__global__ void __launch_bounds__(kThreads) k_sample(const C2<T>* __restrict__ s, uint64_t len,
                                                     const double* __restrict__ prefix, const double* __restrict__ sums,
                                                     uint64_t nblocks, const double* __restrict__ total,
                                                     const double* __restrict__ uni, uint32_t shots,
                                                     uint64_t* __restrict__ out) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t waves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  for (uint64_t shot = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; shot < shots; shot += waves) {
    const double target = uni[shot] * total[0];
    // last block whose prefix <= target among blocks with positive mass
    uint64_t lo = 0, hi = nblocks;  // invariant: answer in [lo, hi)
    while (hi - lo > 1) {
      const uint64_t mid = (lo + hi) >> 1;
      if (prefix[mid] <= target) lo = mid;
      else hi = mid;
    }
    uint64_t blk = lo;
    while (blk > 0 && sums[blk] <= 0.0) blk--;          // never land in an empty block
    while (blk + 1 < nblocks && sums[blk] <= 0.0) blk++;
    const double local = target - prefix[blk];
    const uint64_t base = blk * kSampleBlock;
    const uint32_t per = kSampleBlock / 64;             // amplitudes per lane
    double mine = 0.0;
    for (uint32_t j = 0; j < per; j++) {
      const uint64_t i = base + (uint64_t)lane * per + j;
      if (i < len) {
        const C2<T> a = s[i];
        mine += (double)a.x * (double)a.x + (double)a.y * (double)a.y;
      }
    }
    // inclusive scan over lanes
    double incl = mine;
  #pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const double o = __shfl_up(incl, off, 64);
      if ((int)lane >= off) incl += o;
    }
    const double excl = incl - mine;
    const bool hit = mine > 0.0 && local < incl && local >= excl;
    const uint64_t hits = __ballot(hit);
    const uint64_t pos_lanes = __ballot(mine > 0.0);
    int who;
    if (hits) who = __ffsll((long long)hits) - 1;
    else who = pos_lanes ? 63 - __clzll((long long)pos_lanes) : 0;   // rounding overflow: last positive lane
    if ((int)lane == who) {
      const double rem = hits ? local - excl : 1e308;
      double cum = 0.0;
      uint64_t pick = base + (uint64_t)lane * per;
      uint64_t last_pos = pick;
      bool found = false;
      for (uint32_t j = 0; j < per; j++) {
        const uint64_t i = base + (uint64_t)lane * per + j;
        if (i >= len) break;
        const C2<T> a = s[i];
        const double p = (double)a.x * (double)a.x + (double)a.y * (double)a.y;
        if (p > 0.0) {
          last_pos = i;
          cum += p;
          if (!found && rem < cum) {
            pick = i;
            found = true;
          }
        }
      }
      out[shot] = found ? pick : last_pos;
    }
  }
}

// Capped so that a dispatch stays under 2^32 work-items; kernels that can
// exceed the cap loop grid-stride (k_sample's grid is shots * 64 threads).
// psi *= 1 / sqrt(norm[0] + norm[1]), the FP64 norm a preceding reduce left
// on the device; a non-finite or zero norm leaves the state alone.
template <typename T>
__global__ void __launch_bounds__(kThreads) k_scale_by_norm(C2<T>* __restrict__ s, uint64_t len,
                                                             const double* __restrict__ norm) {
  const double t = norm[0] + norm[1];
  if (norm[2] != 0.0 || !(t > 0.0) || !isfinite(t)) return;
  const double f = 1.0 / sqrt(t);
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i0 < len; i0 += kUnroll * stride) {
    C2<T> a[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; u++) {
      const uint64_t i = i0 + u * stride;
      if (i < len) a[u] = s[i];
    }
#pragma unroll
    for (int u = 0; u < kUnroll; u++) {
      const uint64_t i = i0 + u * stride;
      if (i < len) {
        a[u].x = (T)((double)a[u].x * f);
        a[u].y = (T)((double)a[u].y * f);
        s[i] = a[u];
      }
    }
  }
}

// Per block: Re and Im of sum conj(a_i) b_i, sum |a_i - b_i|^2 (compensated,
// FP64) and max |a_i - b_i|^2; a and b may differ in precision.
template <typename TA, typename TB>
__global__ void __launch_bounds__(kThreads) k_compare(const C2<TA>* __restrict__ a, const C2<TB>* __restrict__ b,
                                                       uint64_t len, double* __restrict__ partials) {
  KSum v[3] = {{0, 0}, {0, 0}, {0, 0}};
  double mx = 0.0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += (uint64_t)gridDim.x * blockDim.x) {
    const double ar = a[i].x, ai = a[i].y, br = b[i].x, bi = b[i].y;
    kadd(v[0], ar * br + ai * bi);
    kadd(v[1], ar * bi - ai * br);
    const double dr = ar - br, di = ai - bi, d2 = dr * dr + di * di;
    kadd(v[2], d2);
    mx = fmax(mx, d2);
  }
  block_reduce<3>(v);
  __shared__ double shm[kThreads / 64];
  for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_down(mx, off, 64));
  if ((threadIdx.x & 63u) == 0) shm[threadIdx.x >> 6] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (uint32_t w = 1; w < blockDim.x / 64; w++) mx = fmax(mx, shm[w]);
    partials[4 * blockIdx.x + 0] = v[0].s + v[0].c;
    partials[4 * blockIdx.x + 1] = v[1].s + v[1].c;
    partials[4 * blockIdx.x + 2] = v[2].s + v[2].c;
    partials[4 * blockIdx.x + 3] = mx;
  }
}

inline uint32_t grid_for(uint64_t items) {
  uint64_t g = (items + kThreads - 1) / kThreads;
  if (g > (1ull << 20)) g = 1ull << 20;
  return (uint32_t)(g == 0 ? 1 : g);
}

}  // namespace

template <typename T>
void launch_reduce_branch(hipStream_t st, const void* state, uint32_t n, int32_t qubit, double* partials,
                          double* out) {
  const uint64_t len = 1ull << n;
  const uint64_t qmask = qubit >= 0 ? (1ull << qubit) : 0ull;
  uint32_t blocks = grid_for((len + kUnroll - 1) / kUnroll);
  if (blocks > kReduceBlocks) blocks = kReduceBlocks;
  hipLaunchKernelGGL((k_reduce_branch<T, 0>), dim3(blocks), dim3(kThreads), 0, st, (const C2<T>*)state, len, qmask,
                     partials);
  hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(kThreads), 0, st, partials, blocks, 3u, out);
  ALEPH_HIP_CHECK(hipGetLastError());
}

template <typename T>
void launch_sum_p2(hipStream_t st, const void* state, uint32_t n, double* partials, double* out) {
  const uint64_t len = 1ull << n;
  uint32_t blocks = grid_for((len + kUnroll - 1) / kUnroll);
  if (blocks > kReduceBlocks) blocks = kReduceBlocks;
  hipLaunchKernelGGL((k_reduce_branch<T, 1>), dim3(blocks), dim3(kThreads), 0, st, (const C2<T>*)state, len, 0ull,
                     partials);
  hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(kThreads), 0, st, partials, blocks, 3u, out);
  ALEPH_HIP_CHECK(hipGetLastError());
}

template <typename T>
void launch_expect_group(hipStream_t st, const void* state, uint32_t n, uint64_t xmask, const uint64_t* zy,
                         const uint8_t* use_im, uint32_t count, double* partials, double* out) {
  if (count == 0 || count > (uint32_t)kGroupTerms) throw Failure{ALEPH_HIP_ERR_INVALID_ARGUMENT, "group size"};
  GroupTerms g{};
  for (uint32_t t = 0; t < count; t++) {
    g.zy[t] = zy[t];
    if (use_im[t]) g.use_im |= 1u << t;
  }
  g.count = count;
  uint32_t hbit = 0;
  if (xmask) hbit = 63 - __builtin_clzll(xmask);
  const uint64_t pairs = xmask ? (1ull << (n - 1)) : (1ull << n);
  uint32_t blocks = grid_for(pairs);
  if (blocks > kReduceBlocks) blocks = kReduceBlocks;
  hipLaunchKernelGGL(k_expect_group<T>, dim3(blocks), dim3(kThreads), 0, st, (const C2<T>*)state, pairs, xmask, hbit,
                     g, partials);
  hipLaunchKernelGGL(k_reduce_columns, dim3(1), dim3(kThreads), 0, st, partials, blocks, (uint32_t)kGroupTerms, out);
  ALEPH_HIP_CHECK(hipGetLastError());
}

template <typename T>
void launch_collapse(hipStream_t st, void* state, uint32_t n, uint32_t qubit, int32_t outcome, double scale) {
  const uint64_t len = 1ull << n;
  hipLaunchKernelGGL(k_collapse<T>, dim3(grid_for(len)), dim3(kThreads), 0, st, (C2<T>*)state, len, 1ull << qubit,
                     outcome, scale);
  ALEPH_HIP_CHECK(hipGetLastError());
}

template <typename T>
void launch_marginal(hipStream_t st, const void* state, uint32_t n, uint32_t k, const uint32_t* qubits, double* out) {
  const uint64_t len = 1ull << n;
  const uint64_t bins = 1ull << k;
  hipLaunchKernelGGL(k_zero, dim3(grid_for(bins)), dim3(kThreads), 0, st, out, bins);
  Qubits64 qs{};
  for (uint32_t j = 0; j < k && j < 64; j++) qs.q[j] = qubits[j];
  uint32_t blocks = grid_for(len);
  if (blocks > kReduceBlocks) blocks = kReduceBlocks;
  const size_t lds = k <= 12 ? sizeof(double) * bins : 0;
  hipLaunchKernelGGL(k_marginal<T>, dim3(blocks), dim3(kThreads), lds, st, (const C2<T>*)state, len, k, qs, out);
  ALEPH_HIP_CHECK(hipGetLastError());
}

template <typename T>
void launch_sample(hipStream_t st, const void* state, uint32_t n, double* block_sums, double* scan_tmp,
                   const double* uniforms, uint32_t shots, uint64_t* out) {
  const uint64_t len = 1ull << n;
  const uint64_t nb = n > kSampleBlockBits ? (1ull << (n - kSampleBlockBits)) : 1ull;
  double* sums = block_sums;       // nb
  double* prefix = block_sums + nb;  // nb
  double* total = scan_tmp;
  hipLaunchKernelGGL(k_block_sums<T>, dim3((uint32_t)nb), dim3(kThreads), 0, st, (const C2<T>*)state, len, sums,
                     prefix);
  hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, st, prefix, nb, total);
  if (shots) {
    const uint64_t threads = (uint64_t)shots * 64;
    hipLaunchKernelGGL(k_sample<T>, dim3(grid_for(threads)), dim3(kThreads), 0, st, (const C2<T>*)state, len, prefix,
                       sums, nb, total, uniforms, shots, out);
  }
  ALEPH_HIP_CHECK(hipGetLastError());
}

template <typename T>
void launch_normalize(hipStream_t st, void* state, uint32_t n, double* partials, double* res) {
  launch_reduce_branch<T>(st, state, n, -1, partials, res);
  const uint64_t len = 1ull << n;
  hipLaunchKernelGGL(k_scale_by_norm<T>, dim3(grid_for((len + kUnroll - 1) / kUnroll)), dim3(kThreads), 0, st,
                     (C2<T>*)state, len,
                     (const double*)res);
  ALEPH_HIP_CHECK(hipGetLastError());
}

template <typename TA, typename TB>
uint32_t launch_compare(hipStream_t st, const void* a, const void* b, uint32_t n, double* partials) {
  const uint64_t len = 1ull << n;
  uint32_t blocks = grid_for(len);
  if (blocks > kReduceBlocks) blocks = kReduceBlocks;
  hipLaunchKernelGGL((k_compare<TA, TB>), dim3(blocks), dim3(kThreads), 0, st, (const C2<TA>*)a, (const C2<TB>*)b,
                     len, partials);
  ALEPH_HIP_CHECK(hipGetLastError());
  return blocks;
}

template uint32_t launch_compare<double, double>(hipStream_t, const void*, const void*, uint32_t, double*);
template uint32_t launch_compare<double, float>(hipStream_t, const void*, const void*, uint32_t, double*);
template uint32_t launch_compare<float, double>(hipStream_t, const void*, const void*, uint32_t, double*);
template uint32_t launch_compare<float, float>(hipStream_t, const void*, const void*, uint32_t, double*);

#define ALEPH_INST(T)                                                                                           \
  template void launch_reduce_branch<T>(hipStream_t, const void*, uint32_t, int32_t, double*, double*);         \
  template void launch_sum_p2<T>(hipStream_t, const void*, uint32_t, double*, double*);                         \
  template void launch_normalize<T>(hipStream_t, void*, uint32_t, double*, double*);                            \
  template void launch_expect_group<T>(hipStream_t, const void*, uint32_t, uint64_t, const uint64_t*,            \
                                       const uint8_t*, uint32_t, double*, double*);                             \
  template void launch_collapse<T>(hipStream_t, void*, uint32_t, uint32_t, int32_t, double);                    \
  template void launch_marginal<T>(hipStream_t, const void*, uint32_t, uint32_t, const uint32_t*, double*);      \
  template void launch_sample<T>(hipStream_t, const void*, uint32_t, double*, double*, const double*, uint32_t,   \
                                 uint64_t*);
ALEPH_INST(double)
ALEPH_INST(float)
#undef ALEPH_INST

}  // namespace aleph
