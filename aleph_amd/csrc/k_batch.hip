// Batched execution of many small fused plans on one gfx950 device.
//
// Every instance owns a 2^n_i-amplitude state inside one pool allocation
// and a fused plan (passes of chunks, see fused_chunk.hpp).  Two executors
// run the same work items (BatchItem: some chunks of one pass of one
// instance):
//  - rounds: one launch per pass index, each workgroup one item; kernel
//    boundaries order the passes of an instance;
//  - persistent: one launch of as many workgroups as stay resident, which
//    claim items from a queue in order and wait (bounded) on the
//    instance's done counter for its earlier passes.  Items are queued so
//    that every dependency of an item sits earlier in the queue, hence is
//    already claimed by a running workgroup and never waits on a later
//    item: the waits end.
// The reference runs such sweeps as one circuit per call
// (crates/aleph-backend/src/lib.rs:211-248, run_with_outcomes).
#include "common.hpp"
#include "fused_chunk.hpp"
#include "kernels.hpp"

namespace aleph {
namespace {

using namespace fused;

constexpr int kBatchFeat = kFeat2Q | kFeatPhase;

template <typename T, int G, int BT, bool NT>
__global__ void __launch_bounds__(BT, 2)
    k_batch_round(C2<T>* __restrict__ pool, const BatchItem* __restrict__ items,
                  const PassDev* __restrict__ passes, const StageDev* __restrict__ stages,
                  const DevOp* __restrict__ ops, const PhaseTerm* __restrict__ terms) {
  extern __shared__ __align__(16) unsigned char smem[];
  const BatchItem it = items[blockIdx.x];
  C2<T>* st = pool + it.state_off;
  const PassDev* pd = passes + it.pass;
  for (uint32_t c = 0; c < it.chunk_count; c++) {
    if (c) __syncthreads();
    fused_chunk<T, G, kBatchFeat, NT, true>(st, pd, stages, ops, terms, it.chunk_begin + c,
                                            reinterpret_cast<C2<T>*>(smem));
  }
}

template <typename T, int G, int BT, bool NT>
__global__ void __launch_bounds__(BT, 2)
    k_batch_persistent(C2<T>* __restrict__ pool, const BatchItem* __restrict__ items, uint32_t n_items,
                       const PassDev* __restrict__ passes, const StageDev* __restrict__ stages,
                       const DevOp* __restrict__ ops, const PhaseTerm* __restrict__ terms, uint32_t* done,
                       BatchCtl* ctl, uint32_t spin_limit) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ uint32_t s_k, s_ok;
  const uint32_t tid = threadIdx.x;
  // The previous item is counted done by thread 0 right before it claims
  // the next one, so every divergent (thread-0) region sits between the
  // loop head and the first barrier: with the count at the end of the body
  // instead, the compiler joins the two thread-0 regions across the back
  // edge and splits the loop so that the other lanes return to the barrier
  // without a new claim and re-read a stale slot forever.  The loop
  // control values go through readfirstlane for the same reason.
  uint32_t prev_inst = 0, prev_count = 0;
  for (;;) {
    if (tid == 0) {
      if (prev_count) __hip_atomic_fetch_add(done + prev_inst, prev_count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      s_k = __hip_atomic_fetch_add(&ctl->head, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    const uint32_t k = __builtin_amdgcn_readfirstlane(s_k);
    if (k >= n_items) break;
    const BatchItem it = items[k];
    if (tid == 0) {
      uint32_t ok = 1;
      if (it.dep_target != 0) {
        uint32_t spins = 0;
        while (__hip_atomic_load(done + it.inst, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < it.dep_target) {
          if (__hip_atomic_load(&ctl->error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0 ||
              ++spins > spin_limit) {
            ok = 0;
            break;
          }
          __builtin_amdgcn_s_sleep(4);
        }
        if (!ok) __hip_atomic_store(&ctl->error, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      s_ok = ok;
    }
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    if (__builtin_amdgcn_readfirstlane(s_ok)) {
      C2<T>* st = pool + it.state_off;
      const PassDev* pd = passes + it.pass;
      for (uint32_t c = 0; c < it.chunk_count; c++) {
        if (c) __syncthreads();
        fused_chunk<T, G, kBatchFeat, NT, true>(st, pd, stages, ops, terms, it.chunk_begin + c,
                                                reinterpret_cast<C2<T>*>(smem));
      }
    }
    // Publish this item's stores device-wide before it is counted done (a
    // skipped item is counted too, so no wait depends on it forever).
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    prev_inst = it.inst;
    prev_count = it.chunk_count;
  }
}

// Every instance's initial state in one write: blockIdx.y (grid-stride)
// picks the instance, the x dimension strides over its amplitudes.
template <typename T>
__global__ void __launch_bounds__(256) k_batch_prep(C2<T>* __restrict__ pool, const BatchPrep* __restrict__ prep,
                                                    uint32_t n_inst) {
  for (uint32_t i = blockIdx.y; i < n_inst; i += gridDim.y) {
    const BatchPrep d = prep[i];
    const uint64_t size = 1ull << d.n;
    C2<T>* st = pool + d.state_off;
    for (uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; x < size;
         x += (uint64_t)gridDim.x * blockDim.x) {
      C2<T> v;
      v.x = (x & ~d.plus) == 0 ? (T)d.amp : (T)0;
      v.y = 0;
      st[x] = v;
    }
  }
}

__global__ void __launch_bounds__(256) k_batch_reset(uint32_t* __restrict__ done, uint32_t n_inst,
                                                     BatchCtl* __restrict__ ctl) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n_inst) done[i] = 0;
  if (i == 0) {
    ctl->head = 0;
    ctl->error = 0;
  }
}

// Sum of |psi_x|^2 * cut(x) over blocks of 4096 amplitudes.  A thread owns
// 16 consecutive amplitudes (qubits 0..3 vary, the rest fixed): edges among
// the fixed qubits are counted once per thread, edges from qubit g < 4 to a
// fixed qubit contribute a_g (set neighbours) when bit g is 0 and
// b_g - a_g when it is 1, and edges inside 0..3 come from the 16-entry
// table ll.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

template <typename T>
__global__ void __launch_bounds__(256) k_batch_cut(const C2<T>* __restrict__ pool, const CutDesc* __restrict__ descs,
                                                   uint32_t n_inst, double* __restrict__ partials) {
  __shared__ uint32_t s_inst;
  __shared__ double s_wave[4];
  const uint32_t b = blockIdx.x;
  if (threadIdx.x == 0) {
    uint32_t lo = 0, hi = n_inst;  // last instance with block_begin <= b
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) / 2;
      if (descs[mid].block_begin <= b) lo = mid;
      else hi = mid;
    }
    s_inst = lo;
  }
  __syncthreads();
  const CutDesc& d = descs[s_inst];
  const uint32_t n = d.n;
  const uint64_t amps = 1ull << n;
  const uint64_t first = ((uint64_t)(b - d.block_begin) << kBatchCutBlockBits) + 16ull * threadIdx.x;
  double acc = 0.0;
  if (first < amps) {
    const uint32_t x = (uint32_t)first;
    uint32_t c0 = 0;
    for (uint32_t u = 4; u < n; u++)
      if ((x >> u) & 1u) c0 += __popc(~x & d.adj[u] & ~0xFu);
    uint32_t a[4], bb[4];
#pragma unroll
    for (int g = 0; g < 4; g++) {
      a[g] = __popc(x & d.adj[g] & ~0xFu);
      bb[g] = __popc(d.adj[g] & ~0xFu);
    }
    const uint64_t ll = d.ll;
    const C2<T>* p = pool + d.state_off + first;
#pragma unroll
    for (int r = 0; r < 16; r++) {
      uint32_t cut = c0 + (uint32_t)((ll >> (4 * r)) & 15u);
#pragma unroll
      for (int g = 0; g < 4; g++) cut += ((r >> g) & 1) ? bb[g] - a[g] : a[g];
      const double re = (double)p[r].x, im = (double)p[r].y;
      acc += (re * re + im * im) * (double)cut;
    }
  }
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) s_wave[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) partials[b] = (s_wave[0] + s_wave[1]) + (s_wave[2] + s_wave[3]);
}

// One workgroup per instance sums its partials (Neumaier per lane).
__global__ void __launch_bounds__(256) k_batch_cut_final(const CutDesc* __restrict__ descs, uint32_t n_inst,
                                                         uint32_t n_blocks, const double* __restrict__ partials,
                                                         double* __restrict__ out) {
  __shared__ double s_wave[4];
  const uint32_t i = blockIdx.x;
  const uint32_t b0 = descs[i].block_begin;
  const uint32_t b1 = i + 1 < n_inst ? descs[i + 1].block_begin : n_blocks;
  double s = 0.0, c = 0.0;
  for (uint32_t b = b0 + threadIdx.x; b < b1; b += blockDim.x) {
    const double v = partials[b];
    const double t = s + v;
    c += fabs(s) >= fabs(v) ? (s - t) + v : (v - t) + s;
    s = t;
  }
  double tot = wave_sum(s + c);
  if ((threadIdx.x & 63) == 0) s_wave[threadIdx.x >> 6] = tot;
  __syncthreads();
  if (threadIdx.x == 0) out[i] = (s_wave[0] + s_wave[1]) + (s_wave[2] + s_wave[3]);
}

// Dynamic LDS up to the CU's 160 KiB less what the kernels declare
// statically (the persistent executor's queue slots).
template <typename K>
void allow_big_lds(K kernel) {
  ALEPH_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - 1024));
}

template <typename T, int G, int BT, bool NT>
void round_bt(hipStream_t s, void* pool, const BatchItem* items, uint32_t n_items, size_t lds, const PassDev* passes,
              const StageDev* stages, const DevOp* ops, const PhaseTerm* terms) {
  hipLaunchKernelGGL((k_batch_round<T, G, BT, NT>), dim3(n_items), dim3(BT), lds, s,
                     reinterpret_cast<C2<T>*>(pool), items, passes, stages, ops, terms);
  ALEPH_HIP_CHECK(hipGetLastError());
}

template <typename T, int G, int BT, bool NT>
void persistent_bt(hipStream_t s, void* pool, const BatchItem* items, uint32_t n_items, size_t lds, uint32_t grid,
                   const PassDev* passes, const StageDev* stages, const DevOp* ops, const PhaseTerm* terms,
                   uint32_t* done, BatchCtl* ctl, uint32_t spin_limit) {
  hipLaunchKernelGGL((k_batch_persistent<T, G, BT, NT>), dim3(grid), dim3(BT), lds, s,
                     reinterpret_cast<C2<T>*>(pool), items, n_items, passes, stages, ops, terms, done, ctl,
                     spin_limit);
  ALEPH_HIP_CHECK(hipGetLastError());
}

template <typename T, int G, int BT, bool NT>
int occupancy_bt(size_t lds) {
  allow_big_lds(&k_batch_persistent<T, G, BT, NT>);
  int blocks = 0;
  ALEPH_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
      &blocks, reinterpret_cast<const void*>(&k_batch_persistent<T, G, BT, NT>), BT, lds));
  return blocks;
}

}  // namespace

// Dispatch on (BT, NT); every pass of a batch uses G = 4 (G = 5 spills here).
#define ALEPH_BATCH_DISPATCH(FN, ...)                                   \
  do {                                                                  \
    if (cfg.threads <= 256) {                                           \
      if (cfg.nontemporal) FN<T, 4, 256, true>(__VA_ARGS__);            \
      else FN<T, 4, 256, false>(__VA_ARGS__);                           \
    } else {                                                            \
      if (cfg.nontemporal) FN<T, 4, 512, true>(__VA_ARGS__);            \
      else FN<T, 4, 512, false>(__VA_ARGS__);                           \
    }                                                                   \
  } while (0)

template <typename T, int G, int BT, bool NT>
void configure_bt() {
  allow_big_lds(&k_batch_round<T, G, BT, NT>);
  allow_big_lds(&k_batch_persistent<T, G, BT, NT>);
}

template <typename T>
void batch_configure(const BatchLaunch& cfg) {
  ALEPH_BATCH_DISPATCH(configure_bt, );
}

template <typename T>
void launch_batch_round(hipStream_t s, const BatchLaunch& cfg, void* pool, const BatchItem* items, uint32_t n_items,
                        const PassDev* passes, const StageDev* stages, const DevOp* ops, const PhaseTerm* terms) {
  if (n_items == 0) return;
  ALEPH_BATCH_DISPATCH(round_bt, s, pool, items, n_items, cfg.lds_bytes, passes, stages, ops, terms);
}

template <typename T>
void launch_batch_persistent(hipStream_t s, const BatchLaunch& cfg, void* pool, const BatchItem* items,
                             uint32_t n_items, uint32_t grid, const PassDev* passes, const StageDev* stages,
                             const DevOp* ops, const PhaseTerm* terms, uint32_t* done, BatchCtl* ctl,
                             uint32_t spin_limit) {
  if (n_items == 0) return;
  ALEPH_BATCH_DISPATCH(persistent_bt, s, pool, items, n_items, cfg.lds_bytes, grid, passes, stages, ops, terms, done,
                       ctl, spin_limit);
}

template <typename T>
int batch_persistent_occupancy(const BatchLaunch& cfg) {
  int r = 0;
  if (cfg.threads <= 256) r = cfg.nontemporal ? occupancy_bt<T, 4, 256, true>(cfg.lds_bytes)
                                              : occupancy_bt<T, 4, 256, false>(cfg.lds_bytes);
  else r = cfg.nontemporal ? occupancy_bt<T, 4, 512, true>(cfg.lds_bytes)
                           : occupancy_bt<T, 4, 512, false>(cfg.lds_bytes);
  return r;
}

template <typename T>
void launch_batch_init(hipStream_t s, void* pool, const BatchPrep* prep, uint32_t n_inst, uint32_t* done,
                       BatchCtl* ctl) {
  if (n_inst) {
    hipLaunchKernelGGL((k_batch_prep<T>), dim3(64, std::min<uint32_t>(n_inst, 65535)), dim3(256), 0, s,
                       reinterpret_cast<C2<T>*>(pool), prep, n_inst);
    ALEPH_HIP_CHECK(hipGetLastError());
  }
  launch_batch_reset(s, done, n_inst, ctl);
}

void launch_batch_reset(hipStream_t s, uint32_t* done, uint32_t n_inst, BatchCtl* ctl) {
  const uint32_t ib = (n_inst + 255) / 256;
  hipLaunchKernelGGL(k_batch_reset, dim3(ib ? ib : 1), dim3(256), 0, s, done, n_inst, ctl);
  ALEPH_HIP_CHECK(hipGetLastError());
}

template <typename T>
void launch_batch_cut(hipStream_t s, const void* pool, const CutDesc* descs, uint32_t n_inst, uint32_t n_blocks,
                      double* partials, double* out) {
  if (n_inst == 0) return;
  hipLaunchKernelGGL((k_batch_cut<T>), dim3(n_blocks), dim3(256), 0, s, reinterpret_cast<const C2<T>*>(pool), descs,
                     n_inst, partials);
  ALEPH_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_batch_cut_final, dim3(n_inst), dim3(256), 0, s, descs, n_inst, n_blocks, partials, out);
  ALEPH_HIP_CHECK(hipGetLastError());
}

#define ALEPH_BATCH_INST(T)                                                                                  \
  template void launch_batch_round<T>(hipStream_t, const BatchLaunch&, void*, const BatchItem*, uint32_t,    \
                                      const PassDev*, const StageDev*, const DevOp*, const PhaseTerm*);      \
  template void launch_batch_persistent<T>(hipStream_t, const BatchLaunch&, void*, const BatchItem*,         \
                                           uint32_t, uint32_t, const PassDev*, const StageDev*,              \
                                           const DevOp*, const PhaseTerm*, uint32_t*, BatchCtl*, uint32_t);  \
  template int batch_persistent_occupancy<T>(const BatchLaunch&);                                            \
  template void batch_configure<T>(const BatchLaunch&);                                                      \
  template void launch_batch_init<T>(hipStream_t, void*, const BatchPrep*, uint32_t, uint32_t*, BatchCtl*);  \
  template void launch_batch_cut<T>(hipStream_t, const void*, const CutDesc*, uint32_t, uint32_t, double*,   \
                                    double*);
ALEPH_BATCH_INST(double)
ALEPH_BATCH_INST(float)

}  // namespace aleph
