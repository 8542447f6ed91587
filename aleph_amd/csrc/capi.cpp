// C ABI of the gfx950 state-vector backend (include/aleph_hip.h).
//
// Every entry point validates on the host before any launch (as
// crates/aleph-cuda/src/common.rs `validate_and_extract` does), converts
// failures into a status code plus structured detail kept on the context,
// and orders all device work on the context's single stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <vector>

#include "batch.hpp"
#include "common.hpp"
#include "engine.hpp"
#include "gates.hpp"
#include "jit.hpp"
#include "kernels.hpp"
#include "noise_host.hpp"

using namespace aleph;

// Device memory of the allocation-free sampler
// (aleph_hip_sample_with_uniforms_noalloc): the block sums and their prefix
// for a state of up to kNoAllocMaxQubits qubits, the scan's total, and up to
// kNoAllocMaxShots uniforms and results: 17 MiB.  One copy per device, held
// until the process ends.  In host code `scratch` only names the symbol; its
// device address comes from hipGetSymbolAddress (noalloc_scratch_address).
namespace aleph {
constexpr uint32_t kNoAllocMaxQubits = 32;
constexpr uint32_t kNoAllocMaxShots = 1u << 16;
constexpr size_t kNoAllocBlocks = size_t(1) << (kNoAllocMaxQubits - kSampleBlockBits);
constexpr size_t kNoAllocSumsOff = 0;                                          // 2 * kNoAllocBlocks doubles
constexpr size_t kNoAllocMiscOff = kNoAllocSumsOff + 2 * kNoAllocBlocks * sizeof(double);  // 64 doubles
constexpr size_t kNoAllocUniOff = kNoAllocMiscOff + 64 * sizeof(double);       // kNoAllocMaxShots doubles
constexpr size_t kNoAllocOutOff = kNoAllocUniOff + kNoAllocMaxShots * sizeof(double);  // kNoAllocMaxShots u64
constexpr size_t kNoAllocBytes = kNoAllocOutOff + kNoAllocMaxShots * sizeof(uint64_t);
}  // namespace aleph

static __device__ __align__(256) unsigned char scratch[aleph::kNoAllocBytes];

struct aleph_hip_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  uint64_t rng[4] = {0, 0, 0, 0};
  uint32_t cap = 0;                 // 0: 32 qubits for F64, 33 for F32
  aleph_hip_error err{};
  double* d_scratch = nullptr;      // reductions + small matrices
  size_t scratch_bytes = 0;
  double* h_pinned = nullptr;       // 64 doubles of pinned readback
  float* d_model = nullptr;         // cost model of the learned planner (kMlpParams floats) or null
  void* noalloc_scratch = nullptr;  // this device's copy of `scratch`
  std::vector<aleph_hip_state*> states;  // live states, detached from the context when it is destroyed
};

struct aleph_hip_state {
  aleph_hip_ctx* ctx = nullptr;
  uint32_t n = 0;
  uint32_t precision = ALEPH_HIP_F64;
  void* d = nullptr;
};

struct aleph_hip_plan {
  std::unique_ptr<Plan> plan;
};

struct aleph_hip_batch {
  std::unique_ptr<Batch> b;
};

struct aleph_hip_model {
  float* d = nullptr;   // kMlpParams floats on the device
  ~aleph_hip_model() {
    if (d) (void)hipFree(d);
  }
};

namespace {

constexpr size_t kScratchBytes = 2u << 20;
// Context scratch layout (doubles): reduction partials from 0 (at most
// kExpectGroup * kReduceBlocks of them), results from kResOffset.
constexpr size_t kResOffset = (size_t)kExpectGroup * kReduceBlocks;
static_assert((kResOffset + 4096) * sizeof(double) <= kScratchBytes, "scratch layout");

// xoshiro256** seeded through SplitMix64.
void rng_seed(aleph_hip_ctx* c, uint64_t seed) {
  for (int i = 0; i < 4; i++) {
    seed += 0x9E3779B97F4A7C15ull;
    uint64_t z = seed;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    c->rng[i] = z ^ (z >> 31);
  }
}

uint64_t rng_next(aleph_hip_ctx* c) {
  uint64_t* s = c->rng;
  auto rotl = [](uint64_t x, int k) { return (x << k) | (x >> (64 - k)); };
  const uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
  s[2] ^= s[0];
  s[3] ^= s[1];
  s[1] ^= s[2];
  s[0] ^= s[3];
  s[2] ^= t;
  s[3] = rotl(s[3], 45);
  return r;
}

double rng_uniform(aleph_hip_ctx* c) { return (double)(rng_next(c) >> 11) * (1.0 / 9007199254740992.0); }

void set_error(aleph_hip_ctx* c, const Failure& f) {
  if (!c) return;
  c->err = aleph_hip_error{};
  c->err.code = f.code;
  c->err.qubit = f.qubit;
  c->err.num_qubits = f.num_qubits;
  c->err.expected = f.expected;
  c->err.got = f.got;
  c->err.value = f.value;
  std::snprintf(c->err.kind, sizeof(c->err.kind), "%s", f.kind.c_str());
  std::snprintf(c->err.message, sizeof(c->err.message), "%s", f.message.c_str());
}

template <typename F>
aleph_hip_status guarded(aleph_hip_ctx* c, F&& body) {
  try {
    if (c) {
      c->err = aleph_hip_error{};
      ALEPH_HIP_CHECK(hipSetDevice(c->device));
    }
    body();
    return ALEPH_HIP_OK;
  } catch (const Failure& f) {
    set_error(c, f);
    return f.code;
  } catch (const std::bad_alloc&) {
    set_error(c, Failure{ALEPH_HIP_ERR_OUT_OF_MEMORY, "host allocation failed"});
    return ALEPH_HIP_ERR_OUT_OF_MEMORY;
  }
}

void need(bool cond, const char* what) {
  if (!cond) {
    Failure f{ALEPH_HIP_ERR_INVALID_ARGUMENT, what};
    f.kind = "argument";
    throw f;
  }
}

void check_state(aleph_hip_ctx* c, aleph_hip_state* st) {
  need(c != nullptr && st != nullptr, "null context or state");
  need(st->ctx == c, "state belongs to another context");
}

void check_qubit(const aleph_hip_state* st, uint32_t q, const char* kind) {
  if (q >= st->n) {
    Failure f = make_failure(ALEPH_HIP_ERR_QUBIT_OUT_OF_RANGE, kind, "qubit out of range");
    f.qubit = q;
    f.num_qubits = st->n;
    throw f;
  }
}

void check_distinct(const uint32_t* qs, uint32_t k, const aleph_hip_state* st, const char* kind) {
  uint64_t seen = 0;
  for (uint32_t i = 0; i < k; i++) {
    check_qubit(st, qs[i], kind);
    if (seen & (1ull << qs[i])) {
      Failure f = make_failure(ALEPH_HIP_ERR_DUPLICATE_QUBIT, kind, "duplicate qubit");
      f.qubit = qs[i];
      throw f;
    }
    seen |= 1ull << qs[i];
  }
}

size_t amp_bytes(const aleph_hip_state* st) { return st->precision == ALEPH_HIP_F64 ? 16 : 8; }

// Device address of `scratch` on the current device `dev`, looked up once per
// device (each device holds its own copy of the array at its own address).
void* noalloc_scratch_address(int dev) {
  static std::mutex mu;
  static std::vector<void*> addr;
  std::lock_guard<std::mutex> lock(mu);
  if ((size_t)dev >= addr.size()) addr.resize((size_t)dev + 1, nullptr);
  if (!addr[dev]) {
    void* p = nullptr;
    ALEPH_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw Failure{ALEPH_HIP_ERR_DEVICE_FAILURE, "no device address for the sampler scratch"};
    addr[dev] = p;
  }
  return addr[dev];
}

// Σ|ψ|² over bit q clear / set, non-finite count (q < 0: everything in [0]).
void reduce(aleph_hip_ctx* c, aleph_hip_state* st, int32_t q, double out[3]) {
  double* partials = c->d_scratch;
  double* res = c->d_scratch + kResOffset;
  if (st->precision == ALEPH_HIP_F64) launch_reduce_branch<double>(c->stream, st->d, st->n, q, partials, res);
  else launch_reduce_branch<float>(c->stream, st->d, st->n, q, partials, res);
  ALEPH_HIP_CHECK(hipMemcpyAsync(c->h_pinned, res, 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  ALEPH_HIP_CHECK(hipStreamSynchronize(c->stream));
  for (int i = 0; i < 3; i++) out[i] = c->h_pinned[i];
}

// FP64 norm of the state, then psi /= sqrt(norm), both on the context stream.
void normalize_async(aleph_hip_ctx* c, aleph_hip_state* st) {
  double* partials = c->d_scratch;
  double* res = c->d_scratch + kResOffset;
  if (st->precision == ALEPH_HIP_F64) launch_normalize<double>(c->stream, st->d, st->n, partials, res);
  else launch_normalize<float>(c->stream, st->d, st->n, partials, res);
}

// validate_state (crates/aleph-sv/src/measure.rs:23-39).
void validate(aleph_hip_ctx* c, aleph_hip_state* st, double* p0, double* p1, int32_t q) {
  double r[3];
  reduce(c, st, q, r);
  const double total = r[0] + r[1];
  const double tol = st->precision == ALEPH_HIP_F32 ? kFp32NormDrift
                                                    : std::sqrt(std::ldexp(1.0, (int)st->n)) * kAmplitudeTol;
  if (r[2] != 0.0) throw make_failure(ALEPH_HIP_ERR_INVALID_STATE, "non-finite amplitude norm²", "invalid state");
  if (!std::isfinite(total) || std::fabs(total - 1.0) > tol)
    throw make_failure(ALEPH_HIP_ERR_INVALID_STATE, "state norm² deviates from 1 beyond drift budget",
                       "invalid state");
  if (p0) *p0 = r[0];
  if (p1) *p1 = r[1];
}

struct DevBuf {
  void* p = nullptr;
  explicit DevBuf(size_t bytes) { ALEPH_HIP_CHECK(hipMalloc(&p, std::max<size_t>(bytes, 16))); }
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};

aleph_hip_status measure_impl(aleph_hip_ctx* c, aleph_hip_state* st, uint32_t qubit, const double* forced_u,
                              int32_t* outcome, int32_t* consumed) {
  return guarded(c, [&] {
    check_state(c, st);
    need(outcome != nullptr, "null outcome");
    check_qubit(st, qubit, "measure");
    double p0, p1;
    validate(c, st, &p0, &p1, (int32_t)qubit);
    p0 = std::min(std::max(p0, 0.0), 1.0);
    p1 = std::min(std::max(p1, 0.0), 1.0);
    const bool one_deg = p1 < 1e-300, zero_deg = p0 < 1e-300;  // measure.rs:12, 77-90
    bool out;
    double p;
    if (consumed) *consumed = 0;
    if (zero_deg && one_deg) {
      Failure f = make_failure(ALEPH_HIP_ERR_DEGENERATE_MEASUREMENT, "measure", "degenerate branch");
      f.qubit = qubit;
      f.value = std::max(p0, p1);
      throw f;
    } else if (zero_deg) {
      out = true;
      p = p1;
    } else if (one_deg) {
      out = false;
      p = p0;
    } else {
      const double u = forced_u ? *forced_u : rng_uniform(c);
      if (consumed) *consumed = 1;
      out = u < p1;
      p = out ? p1 : p0;
    }
    const double scale = 1.0 / std::sqrt(p);
    if (st->precision == ALEPH_HIP_F64) launch_collapse<double>(c->stream, st->d, st->n, qubit, out, scale);
    else launch_collapse<float>(c->stream, st->d, st->n, qubit, out, scale);
    ALEPH_HIP_CHECK(hipStreamSynchronize(c->stream));
    *outcome = out ? 1 : 0;
  });
}

aleph_hip_status sample_impl(aleph_hip_ctx* c, aleph_hip_state* st, const double* uniforms, uint32_t shots,
                             uint64_t* out) {
  return guarded(c, [&] {
    check_state(c, st);
    need(shots == 0 || out != nullptr, "null output");
    validate(c, st, nullptr, nullptr, -1);
    if (shots == 0) return;
    std::vector<double> u(shots);
    for (uint32_t i = 0; i < shots; i++) u[i] = uniforms ? uniforms[i] : rng_uniform(c);
    const uint64_t nb = st->n > kSampleBlockBits ? (1ull << (st->n - kSampleBlockBits)) : 1ull;
    DevBuf blocks(2 * nb * sizeof(double));
    DevBuf misc(64 * sizeof(double));
    DevBuf du(shots * sizeof(double));
    DevBuf dout(shots * sizeof(uint64_t));
    ALEPH_HIP_CHECK(hipMemcpyAsync(du.p, u.data(), shots * sizeof(double), hipMemcpyHostToDevice, c->stream));
    if (st->precision == ALEPH_HIP_F64)
      launch_sample<double>(c->stream, st->d, st->n, (double*)blocks.p, (double*)misc.p, (const double*)du.p, shots,
                            (uint64_t*)dout.p);
    else
      launch_sample<float>(c->stream, st->d, st->n, (double*)blocks.p, (double*)misc.p, (const double*)du.p, shots,
                           (uint64_t*)dout.p);
    ALEPH_HIP_CHECK(hipMemcpyAsync(out, dout.p, shots * sizeof(uint64_t), hipMemcpyDeviceToHost, c->stream));
    ALEPH_HIP_CHECK(hipStreamSynchronize(c->stream));
  });
}

}  // namespace

extern "C" {

aleph_hip_status aleph_hip_device_count(int32_t* count) {
  if (!count) return ALEPH_HIP_ERR_INVALID_ARGUMENT;
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
  *count = c;
  return ALEPH_HIP_OK;
}

aleph_hip_status aleph_hip_ctx_create(int32_t ordinal, uint64_t seed, aleph_hip_ctx** out) {
  if (!out) return ALEPH_HIP_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return ALEPH_HIP_ERR_NO_DEVICE;
  if (ordinal < 0 || ordinal >= count) return ALEPH_HIP_ERR_NO_DEVICE;
  auto* c = new aleph_hip_ctx;
  c->device = ordinal;
  rng_seed(c, seed);
  const aleph_hip_status st = guarded(c, [&] {
    ALEPH_HIP_CHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    ALEPH_HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&c->d_scratch), kScratchBytes));
    c->scratch_bytes = kScratchBytes;
    ALEPH_HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&c->h_pinned), 64 * sizeof(double), hipHostMallocDefault));
    c->noalloc_scratch = noalloc_scratch_address(c->device);
  });
  if (st != ALEPH_HIP_OK) {
    aleph_hip_ctx_destroy(c);
    return st;
  }
  *out = c;
  return ALEPH_HIP_OK;
}

void aleph_hip_ctx_destroy(aleph_hip_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->d_scratch) (void)hipFree(c->d_scratch);
  if (c->d_model) (void)hipFree(c->d_model);
  for (aleph_hip_state* st : c->states) st->ctx = nullptr;  // states may be freed after their context
  if (c->h_pinned) (void)hipHostFree(c->h_pinned);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

aleph_hip_status aleph_hip_ctx_set_qubit_cap(aleph_hip_ctx* c, uint32_t cap) {
  if (!c) return ALEPH_HIP_ERR_INVALID_ARGUMENT;
  c->cap = cap;
  return ALEPH_HIP_OK;
}

aleph_hip_status aleph_hip_ctx_reseed(aleph_hip_ctx* c, uint64_t seed) {
  if (!c) return ALEPH_HIP_ERR_INVALID_ARGUMENT;
  rng_seed(c, seed);
  return ALEPH_HIP_OK;
}

aleph_hip_status aleph_hip_synchronize(aleph_hip_ctx* c) {
  return guarded(c, [&] {
    need(c != nullptr, "null context");
    ALEPH_HIP_CHECK(hipStreamSynchronize(c->stream));
  });
}

aleph_hip_status aleph_hip_last_error(const aleph_hip_ctx* c, aleph_hip_error* out) {
  if (!c || !out) return ALEPH_HIP_ERR_INVALID_ARGUMENT;
  *out = c->err;
  return ALEPH_HIP_OK;
}

aleph_hip_status aleph_hip_ctx_stream(aleph_hip_ctx* c, void** stream) {
  if (!c || !stream) return ALEPH_HIP_ERR_INVALID_ARGUMENT;
  *stream = c->stream;
  return ALEPH_HIP_OK;
}

aleph_hip_status aleph_hip_state_alloc(aleph_hip_ctx* c, uint32_t n, uint32_t precision, aleph_hip_state** out) {
  return guarded(c, [&] {
    need(c != nullptr && out != nullptr, "null argument");
    *out = nullptr;
    need(precision == ALEPH_HIP_F64 || precision == ALEPH_HIP_F32, "precision must be 64 or 32");
    const uint32_t cap = c->cap ? c->cap : (precision == ALEPH_HIP_F64 ? 32u : 33u);
    if (n > cap || n > 40) {
      Failure f = make_failure(ALEPH_HIP_ERR_TOO_MANY_QUBITS, "allocate", "too many qubits");
      f.expected = std::min<uint32_t>(cap, 40);
      f.got = n;
      throw f;
    }
    const size_t bytes = (precision == ALEPH_HIP_F64 ? 16ull : 8ull) << n;
    size_t free_b = 0, total_b = 0;
    ALEPH_HIP_CHECK(hipMemGetInfo(&free_b, &total_b));
    if (bytes > free_b) {
      Failure f{ALEPH_HIP_ERR_OUT_OF_MEMORY, "state does not fit in free device memory"};
      f.kind = "allocate";
      throw f;
    }
    auto st = std::make_unique<aleph_hip_state>();
    st->ctx = c;
    st->n = n;
    st->precision = precision;
    ALEPH_HIP_CHECK(hipMalloc(&st->d, bytes));
    if (precision == ALEPH_HIP_F64) launch_init_basis<double>(c->stream, st->d, n);
    else launch_init_basis<float>(c->stream, st->d, n);
    ALEPH_HIP_CHECK(hipStreamSynchronize(c->stream));
    c->states.push_back(st.get());
    *out = st.release();
  });
}

void aleph_hip_state_free(aleph_hip_state* st) {
  if (!st) return;
  if (st->ctx) {
    (void)hipSetDevice(st->ctx->device);
    (void)hipStreamSynchronize(st->ctx->stream);
    auto& v = st->ctx->states;
    v.erase(std::remove(v.begin(), v.end(), st), v.end());
  }
  if (st->d) (void)hipFree(st->d);
  delete st;
}

aleph_hip_status aleph_hip_state_reset(aleph_hip_ctx* c, aleph_hip_state* st) {
  return guarded(c, [&] {
    check_state(c, st);
    if (st->precision == ALEPH_HIP_F64) launch_init_basis<double>(c->stream, st->d, st->n);
    else launch_init_basis<float>(c->stream, st->d, st->n);
    ALEPH_HIP_CHECK(hipStreamSynchronize(c->stream));
  });
}

aleph_hip_status aleph_hip_state_load(aleph_hip_ctx* c, aleph_hip_state* st, const double* amps) {
  return guarded(c, [&] {
    check_state(c, st);
    need(amps != nullptr, "null amplitudes");
    // Launches on the context stream (a non-blocking stream) may still be
    // writing the state: finish them before the copy overwrites it.
    ALEPH_HIP_CHECK(hipStreamSynchronize(c->stream));
    const size_t len = 1ull << st->n;
    if (st->precision == ALEPH_HIP_F64) {
      ALEPH_HIP_CHECK(hipMemcpy(st->d, amps, len * 16, hipMemcpyHostToDevice));
    } else {
      std::vector<float> f(2 * len);
      for (size_t i = 0; i < 2 * len; i++) f[i] = (float)amps[i];
      ALEPH_HIP_CHECK(hipMemcpy(st->d, f.data(), len * 8, hipMemcpyHostToDevice));
    }
  });
}

aleph_hip_status aleph_hip_state_read(aleph_hip_ctx* c, aleph_hip_state* st, double* out) {
  return guarded(c, [&] {
    check_state(c, st);
    need(out != nullptr, "null output");
    ALEPH_HIP_CHECK(hipStreamSynchronize(c->stream));
    const size_t len = 1ull << st->n;
    if (st->precision == ALEPH_HIP_F64) {
      ALEPH_HIP_CHECK(hipMemcpy(out, st->d, len * 16, hipMemcpyDeviceToHost));
    } else {
      std::vector<float> f(2 * len);
      ALEPH_HIP_CHECK(hipMemcpy(f.data(), st->d, len * 8, hipMemcpyDeviceToHost));
      for (size_t i = 0; i < 2 * len; i++) out[i] = f[i];
    }
  });
}

aleph_hip_status aleph_hip_state_load_device(aleph_hip_ctx* c, aleph_hip_state* st, const void* dptr) {
  return guarded(c, [&] {
    check_state(c, st);
    need(dptr != nullptr, "null pointer");
    ALEPH_HIP_CHECK(hipMemcpyAsync(st->d, dptr, amp_bytes(st) << st->n, hipMemcpyDeviceToDevice, c->stream));
    ALEPH_HIP_CHECK(hipStreamSynchronize(c->stream));
  });
}

aleph_hip_status aleph_hip_state_read_device(aleph_hip_ctx* c, aleph_hip_state* st, void* dptr) {
  return guarded(c, [&] {
    check_state(c, st);
    need(dptr != nullptr, "null pointer");
    ALEPH_HIP_CHECK(hipMemcpyAsync(dptr, st->d, amp_bytes(st) << st->n, hipMemcpyDeviceToDevice, c->stream));
    ALEPH_HIP_CHECK(hipStreamSynchronize(c->stream));
  });
}

aleph_hip_status aleph_hip_state_info(const aleph_hip_state* st, uint32_t* n, uint32_t* precision, void** dptr) {
  if (!st) return ALEPH_HIP_ERR_INVALID_ARGUMENT;
  if (n) *n = st->n;
  if (precision) *precision = st->precision;
  if (dptr) *dptr = st->d;
  return ALEPH_HIP_OK;
}

aleph_hip_status aleph_hip_apply_gate(aleph_hip_ctx* c, aleph_hip_state* st, const aleph_hip_gate* g) {
  return guarded(c, [&] {
    check_state(c, st);
    need(g != nullptr, "null gate");
    need(g->n_qubits == 0 || g->qubits != nullptr, "null qubits");
    need(g->n_controls == 0 || g->controls != nullptr, "null controls");
    const DenseGate d = validate_gate(*g, st->n);
    apply_dense_gate(c->stream, st->d, st->n, st->precision, d, c->d_scratch);
  });
}

aleph_hip_status aleph_hip_apply_diagonal_phase(aleph_hip_ctx* c, aleph_hip_state* st, uint32_t n_terms,
                                                const double* angles, const uint64_t* conds, const uint32_t* offsets) {
  return guarded(c, [&] {
    check_state(c, st);
    need(n_terms == 0 || (angles && offsets), "null term arrays");
    if (n_terms == 0) return;
    const uint32_t nc = offsets[n_terms];
    need(offsets[0] == 0 && (nc == 0 || conds != nullptr), "bad CSR offsets");
    for (uint32_t t = 0; t < n_terms; t++) {
      need(offsets[t] <= offsets[t + 1], "non-monotone CSR offsets");
      if (!std::isfinite(angles[t])) throw make_failure(ALEPH_HIP_ERR_NON_FINITE_PARAM, "DiagonalPhase", "angle");
    }
    for (uint32_t i = 0; i < nc; i++)
      if (st->n < 64 && (conds[i] >> st->n) != 0) {
        Failure f = make_failure(ALEPH_HIP_ERR_QUBIT_OUT_OF_RANGE, "DiagonalPhase", "mask beyond state");
        f.qubit = 63 - __builtin_clzll(conds[i]);
        f.num_qubits = st->n;
        throw f;
      }
    DevBuf da(n_terms * sizeof(double)), dc((size_t)nc * sizeof(uint64_t)), dof((n_terms + 1) * sizeof(uint32_t));
    ALEPH_HIP_CHECK(hipMemcpyAsync(da.p, angles, n_terms * sizeof(double), hipMemcpyHostToDevice, c->stream));
    if (nc) ALEPH_HIP_CHECK(hipMemcpyAsync(dc.p, conds, nc * sizeof(uint64_t), hipMemcpyHostToDevice, c->stream));
    ALEPH_HIP_CHECK(hipMemcpyAsync(dof.p, offsets, (n_terms + 1) * sizeof(uint32_t), hipMemcpyHostToDevice, c->stream));
    if (st->precision == ALEPH_HIP_F64)
      launch_phase_poly<double>(c->stream, st->d, st->n, n_terms, (const double*)da.p, (const uint64_t*)dc.p,
                                (const uint32_t*)dof.p);
    else
      launch_phase_poly<float>(c->stream, st->d, st->n, n_terms, (const double*)da.p, (const uint64_t*)dc.p,
                               (const uint32_t*)dof.p);
    ALEPH_HIP_CHECK(hipStreamSynchronize(c->stream));
  });
}

aleph_hip_status aleph_hip_unpermute(aleph_hip_ctx* c, aleph_hip_state* st, const uint32_t* perm) {
  return guarded(c, [&] {
    check_state(c, st);
    need(perm != nullptr, "null permutation");
    uint64_t seen = 0;
    for (uint32_t l = 0; l < st->n; l++) {
      need(perm[l] < st->n && !(seen & (1ull << perm[l])), "not a permutation of 0..n-1");
      seen |= 1ull << perm[l];
    }
    void* dst = nullptr;
    ALEPH_HIP_CHECK(hipMalloc(&dst, amp_bytes(st) << st->n));
    if (st->precision == ALEPH_HIP_F64) launch_permute<double>(c->stream, st->d, dst, st->n, perm);
    else launch_permute<float>(c->stream, st->d, dst, st->n, perm);
    ALEPH_HIP_CHECK(hipStreamSynchronize(c->stream));
    (void)hipFree(st->d);
    st->d = dst;
  });
}

aleph_hip_status aleph_hip_apply_1q(aleph_hip_ctx* c, aleph_hip_state* st, uint32_t target, uint64_t cm,
                                    const double* m) {
  return guarded(c, [&] {
    check_state(c, st);
    need(m != nullptr, "null matrix");
    check_qubit(st, target, "apply_1q");
    need(st->n == 64 || (cm >> st->n) == 0, "control mask beyond state");
    need(!(cm & (1ull << target)), "control overlaps target");
    if (st->precision == ALEPH_HIP_F64) launch_1q<double>(c->stream, st->d, st->n, target, cm, m);
    else launch_1q<float>(c->stream, st->d, st->n, target, cm, m);
  });
}

aleph_hip_status aleph_hip_apply_kq(aleph_hip_ctx* c, aleph_hip_state* st, uint32_t k, const uint32_t* qubits,
                                    uint64_t cm, const double* m) {
  return guarded(c, [&] {
    check_state(c, st);
    need(qubits != nullptr && m != nullptr, "null argument");
    need(k >= 1 && k <= kMaxFusedK && k <= st->n, "k must be in [1, 6] and <= n");
    check_distinct(qubits, k, st, "apply_kq");
    need(st->n == 64 || (cm >> st->n) == 0, "control mask beyond state");
    for (uint32_t j = 0; j < k; j++) need(!(cm & (1ull << qubits[j])), "control overlaps target");
    const size_t dim = 1ull << k;
    std::vector<float> f;  // outlives the copy (synchronised below)
    if (st->precision == ALEPH_HIP_F64) {
      ALEPH_HIP_CHECK(hipMemcpyAsync(c->d_scratch, m, 2 * dim * dim * sizeof(double), hipMemcpyHostToDevice, c->stream));
      launch_kq<double>(c->stream, st->d, st->n, k, qubits, cm, c->d_scratch);
    } else {
      f.assign(m, m + 2 * dim * dim);
      ALEPH_HIP_CHECK(hipMemcpyAsync(c->d_scratch, f.data(), f.size() * sizeof(float), hipMemcpyHostToDevice, c->stream));
      launch_kq<float>(c->stream, st->d, st->n, k, qubits, cm, c->d_scratch);
    }
    ALEPH_HIP_CHECK(hipStreamSynchronize(c->stream));
  });
}

aleph_hip_status aleph_hip_apply_diag(aleph_hip_ctx* c, aleph_hip_state* st, uint32_t k, const uint32_t* qubits,
                                      uint64_t cm, const double* d) {
  return guarded(c, [&] {
    check_state(c, st);
    need(qubits != nullptr && d != nullptr, "null argument");
    need(k >= 1 && k <= 16 && k <= st->n, "k must be in [1, 16] and <= n");
    check_distinct(qubits, k, st, "apply_diag");
    need(st->n == 64 || (cm >> st->n) == 0, "control mask beyond state");
    const size_t bytes = (2ull << k) * sizeof(double);
    need(bytes <= kScratchBytes, "diagonal too large");
    ALEPH_HIP_CHECK(hipMemcpyAsync(c->d_scratch, d, bytes, hipMemcpyHostToDevice, c->stream));
    if (st->precision == ALEPH_HIP_F64) launch_diag<double>(c->stream, st->d, st->n, k, qubits, cm, c->d_scratch);
    else launch_diag<float>(c->stream, st->d, st->n, k, qubits, cm, c->d_scratch);
    ALEPH_HIP_CHECK(hipStreamSynchronize(c->stream));
  });
}

aleph_hip_status aleph_hip_apply_cnot(aleph_hip_ctx* c, aleph_hip_state* st, uint32_t control, uint32_t target) {
  return guarded(c, [&] {
    check_state(c, st);
    const uint32_t qs[2] = {control, target};
    check_distinct(qs, 2, st, "apply_cnot");
    const double x[8] = {0, 0, 1, 0, 1, 0, 0, 0};
    if (st->precision == ALEPH_HIP_F64) launch_1q<double>(c->stream, st->d, st->n, target, 1ull << control, x);
    else launch_1q<float>(c->stream, st->d, st->n, target, 1ull << control, x);
  });
}

aleph_hip_status aleph_hip_apply_swap(aleph_hip_ctx* c, aleph_hip_state* st, uint32_t a, uint32_t b) {
  return guarded(c, [&] {
    check_state(c, st);
    const uint32_t qs[2] = {a, b};
    check_distinct(qs, 2, st, "apply_swap");
    if (st->precision == ALEPH_HIP_F64) launch_swap<double>(c->stream, st->d, st->n, a, b, 0);
    else launch_swap<float>(c->stream, st->d, st->n, a, b, 0);
  });
}

namespace {
// Candidate search of the learned planner: every combination of chunk
// qubits, pass budget, 4x4 fusion and non-temporal streaming is planned by
// the rule planner, plus dense-block plans of width 3..6; when the caller
// allows generated kernels (ALEPH_HIP_PLAN_JIT) each candidate is also
// priced as the compiled op-loop kernel.  The cost model predicts every
// fused launch (ln microseconds); standalone sweeps are priced at their
// measured rates.  The cheapest candidate replaces the rule planner's own
// plan only when it is predicted at least kLearnedMargin faster, so that
// prediction noise cannot trade a good rule plan for a near-tie.
constexpr double kLearnedMargin = 0.01;

std::unique_ptr<Plan> plan_learned(aleph_hip_ctx* c, const std::vector<ProgInstr>& prog, uint32_t n,
                                   uint32_t precision, PlanOptions base) {
  const auto t0 = std::chrono::steady_clock::now();
  struct Cand {
    uint32_t m, budget;
    bool f2q, nt;
    uint32_t kq;
    std::unique_ptr<Plan> p;
    double us[2] = {0.0, 0.0};  // predicted total: [0] compiled op loop, [1] generated kernels
  };
  std::vector<Cand> cands;
  for (uint32_t m : {10u, 11u, 12u}) {
    if (m > n && m != 10u) continue;
    for (uint32_t budget : {12u, 24u, 48u}) {
      for (bool f2q : {true, false}) {
        for (bool nt : {false, true}) {
          PlanOptions po = base;
          po.planner = ALEPH_HIP_PLANNER_RULE;
          po.local_qubits = std::min(m, n);
          po.max_ops = budget;
          po.fuse_2q = f2q;
          po.nontemporal = nt;
          po.kq_width = 0;
          po.reg_bits = 0;
          cands.push_back(Cand{m, budget, f2q, nt, 0u, build_plan(prog, n, precision, po)});
        }
      }
    }
  }
  // The rule planner's own choice (12-qubit chunks, budget 24, 4x4 fusion).
  size_t rule = 0;
  for (size_t k = 0; k < cands.size(); k++)
    if (std::min(cands[k].m, n) == std::min(12u, n) && cands[k].budget == 24u && cands[k].f2q && !cands[k].nt) {
      rule = k;
      break;
    }
  for (uint32_t kq = 3; kq <= (uint32_t)kMaxFusedK; kq++) {
    if (n < kq + 4) continue;
    PlanOptions po = base;
    po.planner = ALEPH_HIP_PLANNER_RULE;
    po.kq_width = kq;
    po.reg_bits = 0;
    cands.push_back(Cand{std::min<uint32_t>(12u, n), 24u, true, false, kq, build_plan(prog, n, precision, po)});
  }
  const int variants = base.jit ? 2 : 1;
  std::vector<float> rows;
  std::vector<uint32_t> owner;
  for (uint32_t k = 0; k < cands.size(); k++) {
    std::vector<uint32_t> launch;
    const size_t before = rows.size() / kPlanFeatures;
    plan_features(*cands[k].p, rows, launch);
    owner.insert(owner.end(), rows.size() / kPlanFeatures - before, k);
    cands[k].us[0] = cands[k].us[1] = plan_standalone_us(*cands[k].p);
  }
  const uint32_t R = (uint32_t)owner.size();
  if (R) {
    // Rows [0, R): the variant each candidate was built with (base.jit);
    // rows [R, 2R): the same launches as the compiled op loop.
    if (variants == 2) {
      rows.resize(2 * rows.size());
      std::copy(rows.begin(), rows.begin() + (size_t)R * kPlanFeatures, rows.begin() + (size_t)R * kPlanFeatures);
      for (uint32_t r = 0; r < R; r++) rows[((size_t)R + r) * kPlanFeatures] = 0.f;
    }
    const uint32_t total = R * (uint32_t)variants;
    DevBuf dx(rows.size() * sizeof(float));
    DevBuf dy(total * sizeof(float));
    ALEPH_HIP_CHECK(hipMemcpyAsync(dx.p, rows.data(), rows.size() * sizeof(float), hipMemcpyHostToDevice, c->stream));
    launch_mlp_forward(c->stream, c->d_model, (const float*)dx.p, total, (float*)dy.p);
    std::vector<float> pred(total);
    ALEPH_HIP_CHECK(hipMemcpyAsync(pred.data(), dy.p, total * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    ALEPH_HIP_CHECK(hipStreamSynchronize(c->stream));
    const double scale = std::ldexp(1.0, (int)n - kCostRefQubits);
    for (uint32_t r = 0; r < R; r++) {
      const double t = scale * std::exp((double)std::min(pred[r], 40.f));
      cands[owner[r]].us[base.jit ? 1 : 0] += t;
      if (variants == 2) cands[owner[r]].us[0] += scale * std::exp((double)std::min(pred[R + r], 40.f));
    }
  }
  const int own = base.jit ? 1 : 0;
  size_t best = rule;
  int best_v = own;
  for (size_t k = 0; k < cands.size(); k++)
    for (int v = 0; v < 2; v++) {
      if (variants == 1 && v != own) continue;
      if (cands[k].us[v] < cands[best].us[best_v]) {
        best = k;
        best_v = v;
      }
    }
  const double rule_us = cands[rule].us[own];
  if (cands[best].us[best_v] > (1.0 - kLearnedMargin) * rule_us) {
    best = rule;
    best_v = own;
  }
  std::unique_ptr<Plan> p = std::move(cands[best].p);
  p->opts.planner = ALEPH_HIP_PLANNER_LEARNED;
  p->opts.jit = best_v == 1;
  aleph_hip_plan_stats& st = p->stats;
  st.planner_used = ALEPH_HIP_PLANNER_LEARNED;
  st.chosen_local_qubits = std::min(cands[best].m, n);
  st.chosen_max_ops = cands[best].budget;
  st.chosen_flags = (cands[best].f2q ? 0u : (uint32_t)ALEPH_HIP_PLAN_NO_FUSE_2Q) |
                    (cands[best].nt ? (uint32_t)ALEPH_HIP_PLAN_NONTEMPORAL : 0u) |
                    (best_v == 1 ? (uint32_t)ALEPH_HIP_PLAN_JIT : 0u);
  st.chosen_kq_width = cands[best].kq;
  st.chosen_rule = best == rule ? 1u : 0u;
  st.n_candidates = (uint32_t)(cands.size() * (size_t)variants);
  st.predicted_us = cands[best].us[best_v];
  st.rule_predicted_us = rule_us;
  st.plan_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return p;
}

std::unique_ptr<aleph_hip_plan> make_plan(aleph_hip_ctx* c, uint32_t n, uint32_t precision, const uint32_t* words,
                                          size_t n_words, const double* reals, size_t n_reals,
                                          const aleph_hip_plan_opts* opts) {
    need(precision == ALEPH_HIP_F64 || precision == ALEPH_HIP_F32, "precision must be 64 or 32");
    need(n >= 1 && n <= 40, "qubit count must be in [1, 40]");
    need(n_words == 0 || words != nullptr, "null words");
    need(n_reals == 0 || reals != nullptr, "null reals");
    PlanOptions po;
    if (opts) {
      po.planner = opts->planner;
      po.local_qubits = opts->local_qubits;
      po.max_ops = opts->max_ops_per_pass;
      po.coalesce_bits = opts->coalesce_bits;
      po.fuse_1q = opts->fuse_1q != 0;
      po.reg_bits = opts->reg_bits;
      po.nontemporal = (opts->flags & ALEPH_HIP_PLAN_NONTEMPORAL) != 0;
      po.jit = (opts->flags & ALEPH_HIP_PLAN_JIT) != 0;
      po.fuse_2q = (opts->flags & ALEPH_HIP_PLAN_NO_FUSE_2Q) == 0;
      po.fuse_diag = (opts->flags & ALEPH_HIP_PLAN_NO_FUSE_DIAG) == 0;
      po.kq_width = opts->kq_width;
      po.renormalize = (opts->flags & ALEPH_HIP_PLAN_RENORMALIZE) != 0;
      po.generic_phase = (opts->flags & ALEPH_HIP_PLAN_GENERIC_PHASE) != 0;
    }
    need(po.kq_width == 0 || (po.kq_width >= 3 && po.kq_width <= kMaxFusedK), "kq_width must be 0 or in [3, 6]");
    need(po.reg_bits == 0 || po.reg_bits == 4 || po.reg_bits == 5, "reg_bits must be 4 or 5");
    need(po.planner <= ALEPH_HIP_PLANNER_LEARNED, "unknown planner");
    need(po.local_qubits == 0 || (po.local_qubits >= (po.reg_bits ? po.reg_bits : (uint32_t)kRegBits) &&
                                  po.local_qubits <= (uint32_t)kMaxLocalBits),
         "local_qubits must be in [reg_bits, 13]");
    const std::vector<ProgInstr> prog = decode_program(words, n_words, reals, n_reals, n);
    auto h = std::make_unique<aleph_hip_plan>();
    if (po.planner == ALEPH_HIP_PLANNER_LEARNED && c != nullptr && c->d_model != nullptr &&
        n >= (po.reg_bits ? po.reg_bits : (uint32_t)kRegBits)) {
      h->plan = plan_learned(c, prog, n, precision, po);
    } else {
      h->plan = build_plan(prog, n, precision, po);
      h->plan->stats.planner_used =
          po.planner == ALEPH_HIP_PLANNER_UNFUSED ? ALEPH_HIP_PLANNER_UNFUSED : ALEPH_HIP_PLANNER_RULE;
    }
    return h;
}
}  // namespace

aleph_hip_status aleph_hip_plan_create(aleph_hip_ctx* c, uint32_t n, uint32_t precision, const uint32_t* words,
                                       size_t n_words, const double* reals, size_t n_reals,
                                       const aleph_hip_plan_opts* opts, aleph_hip_plan** out) {
  return guarded(c, [&] {
    need(c != nullptr && out != nullptr, "null argument");
    *out = nullptr;
    auto h = make_plan(c, n, precision, words, n_words, reals, n_reals, opts);
    h->plan->upload(c->stream);
    if (h->plan->opts.jit) jit_compile_plan(*h->plan);
    *out = h.release();
  });
}

aleph_hip_status aleph_hip_plan_build_host(uint32_t n, uint32_t precision, const uint32_t* words, size_t n_words,
                                           const double* reals, size_t n_reals, const aleph_hip_plan_opts* opts,
                                           aleph_hip_plan** out, aleph_hip_error* err) {
  if (!out) return ALEPH_HIP_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  try {
    *out = make_plan(nullptr, n, precision, words, n_words, reals, n_reals, opts).release();
    return ALEPH_HIP_OK;
  } catch (const Failure& f) {
    if (err) {
      aleph_hip_ctx tmp;
      set_error(&tmp, f);
      *err = tmp.err;
    }
    return f.code;
  }
}

aleph_hip_status aleph_hip_plan_view_get(const aleph_hip_plan* plan, aleph_hip_plan_view* out) {
  if (!plan || !plan->plan || !out) return ALEPH_HIP_ERR_INVALID_ARGUMENT;
  const Plan& p = *plan->plan;
  out->steps = p.steps.data();
  out->n_steps = (uint32_t)p.steps.size();
  out->step_bytes = sizeof(Step);
  out->passes = p.passes.data();
  out->n_passes = (uint32_t)p.passes.size();
  out->pass_bytes = sizeof(PassDev);
  out->stages = p.stages.data();
  out->n_stages = (uint32_t)p.stages.size();
  out->stage_bytes = sizeof(StageDev);
  out->ops = p.ops.data();
  out->n_ops = (uint32_t)p.ops.size();
  out->op_bytes = sizeof(DevOp);
  out->terms = p.terms.data();
  out->n_terms = (uint32_t)p.terms.size();
  out->term_bytes = sizeof(PhaseTerm);
  out->payload = p.payload.data();
  out->payload_bytes = p.payload.size();
  return ALEPH_HIP_OK;
}

aleph_hip_status aleph_hip_plan_execute(aleph_hip_ctx* c, aleph_hip_plan* plan, aleph_hip_state* st) {
  return guarded(c, [&] {
    check_state(c, st);
    need(plan != nullptr && plan->plan != nullptr, "null plan");
    need(plan->plan->uploaded, "plan was built on the host only");
    need(plan->plan->n == st->n && plan->plan->precision == st->precision, "plan built for another state shape");
    plan->plan->execute(c->stream, st->d);
    if (plan->plan->opts.renormalize) normalize_async(c, st);
  });
}

aleph_hip_status aleph_hip_plan_execute_timed(aleph_hip_ctx* c, aleph_hip_plan* plan, aleph_hip_state* st,
                                             double* ms, uint32_t* kinds) {
  return guarded(c, [&] {
    check_state(c, st);
    need(plan != nullptr && plan->plan != nullptr, "null plan");
    need(ms != nullptr, "null timing output");
    const Plan& p = *plan->plan;
    need(p.uploaded, "plan was built on the host only");
    need(p.n == st->n && p.precision == st->precision, "plan built for another state shape");
    const size_t ns = p.steps.size();
    std::vector<hipEvent_t> ev(ns + 1, nullptr);
    struct Guard {
      std::vector<hipEvent_t>& e;
      ~Guard() {
        for (hipEvent_t x : e)
          if (x) (void)hipEventDestroy(x);
      }
    } guard{ev};
    for (auto& e : ev) ALEPH_HIP_CHECK(hipEventCreate(&e));
    p.execute_with_events(c->stream, st->d, ev.data());
    if (p.opts.renormalize) normalize_async(c, st);
    ALEPH_HIP_CHECK(hipEventSynchronize(ev[ns]));
    for (size_t i = 0; i < ns; i++) {
      float t = 0.f;
      ALEPH_HIP_CHECK(hipEventElapsedTime(&t, ev[i], ev[i + 1]));
      ms[i] = t;
      if (kinds) kinds[i] = (uint32_t)p.steps[i].kind;
    }
  });
}

aleph_hip_status aleph_hip_plan_stats_get(const aleph_hip_plan* plan, aleph_hip_plan_stats* out) {
  if (!plan || !plan->plan || !out) return ALEPH_HIP_ERR_INVALID_ARGUMENT;
  *out = plan->plan->stats;
  return ALEPH_HIP_OK;
}

void aleph_hip_plan_free(aleph_hip_plan* plan) { delete plan; }

aleph_hip_status aleph_hip_run_program(aleph_hip_ctx* c, aleph_hip_state* st, const uint32_t* words, size_t n_words,
                                       const double* reals, size_t n_reals, const aleph_hip_plan_opts* opts,
                                       aleph_hip_plan_stats* stats) {
  if (!c || !st) return ALEPH_HIP_ERR_INVALID_ARGUMENT;
  aleph_hip_plan* plan = nullptr;
  aleph_hip_status s = aleph_hip_plan_create(c, st->n, st->precision, words, n_words, reals, n_reals, opts, &plan);
  if (s != ALEPH_HIP_OK) return s;
  s = aleph_hip_plan_execute(c, plan, st);
  if (s == ALEPH_HIP_OK) s = aleph_hip_synchronize(c);
  if (s == ALEPH_HIP_OK && stats) aleph_hip_plan_stats_get(plan, stats);
  aleph_hip_plan_free(plan);
  return s;
}

aleph_hip_status aleph_hip_norm_sqr(aleph_hip_ctx* c, aleph_hip_state* st, double* out) {
  return guarded(c, [&] {
    check_state(c, st);
    need(out != nullptr, "null output");
    double r[3];
    reduce(c, st, -1, r);
    *out = r[2] != 0.0 ? NAN : r[0] + r[1];
  });
}

aleph_hip_status aleph_hip_state_normalize(aleph_hip_ctx* c, aleph_hip_state* st, double* norm_before) {
  return guarded(c, [&] {
    check_state(c, st);
    normalize_async(c, st);
    if (norm_before) {
      ALEPH_HIP_CHECK(hipMemcpyAsync(c->h_pinned, c->d_scratch + kResOffset, 3 * sizeof(double),
                                     hipMemcpyDeviceToHost, c->stream));
      ALEPH_HIP_CHECK(hipStreamSynchronize(c->stream));
      *norm_before = c->h_pinned[2] != 0.0 ? NAN : c->h_pinned[0] + c->h_pinned[1];
    }
  });
}

aleph_hip_status aleph_hip_state_compare(aleph_hip_ctx* c, aleph_hip_state* a, aleph_hip_state* b, double* out) {
  return guarded(c, [&] {
    check_state(c, a);
    check_state(c, b);
    need(out != nullptr, "null output");
    need(a->n == b->n, "states differ in size");
    double* partials = c->d_scratch;
    uint32_t blocks = 0;
    const bool a64 = a->precision == ALEPH_HIP_F64, b64 = b->precision == ALEPH_HIP_F64;
    if (a64 && b64) blocks = launch_compare<double, double>(c->stream, a->d, b->d, a->n, partials);
    else if (a64) blocks = launch_compare<double, float>(c->stream, a->d, b->d, a->n, partials);
    else if (b64) blocks = launch_compare<float, double>(c->stream, a->d, b->d, a->n, partials);
    else blocks = launch_compare<float, float>(c->stream, a->d, b->d, a->n, partials);
    std::vector<double> h(4 * (size_t)blocks);
    ALEPH_HIP_CHECK(hipMemcpyAsync(h.data(), partials, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    ALEPH_HIP_CHECK(hipStreamSynchronize(c->stream));
    double re = 0, im = 0, d2 = 0, mx = 0;
    for (uint32_t k = 0; k < blocks; k++) {
      re += h[4 * k];
      im += h[4 * k + 1];
      d2 += h[4 * k + 2];
      mx = std::max(mx, h[4 * k + 3]);
    }
    out[0] = re;
    out[1] = im;
    out[2] = d2;
    out[3] = std::sqrt(mx);
  });
}

aleph_hip_status aleph_hip_validate(aleph_hip_ctx* c, aleph_hip_state* st) {
  return guarded(c, [&] {
    check_state(c, st);
    validate(c, st, nullptr, nullptr, -1);
  });
}

aleph_hip_status aleph_hip_measure(aleph_hip_ctx* c, aleph_hip_state* st, uint32_t qubit, int32_t* outcome) {
  return measure_impl(c, st, qubit, nullptr, outcome, nullptr);
}

aleph_hip_status aleph_hip_measure_with_uniform(aleph_hip_ctx* c, aleph_hip_state* st, uint32_t qubit, double u,
                                                int32_t* outcome, int32_t* consumed) {
  return measure_impl(c, st, qubit, &u, outcome, consumed);
}

aleph_hip_status aleph_hip_sample(aleph_hip_ctx* c, aleph_hip_state* st, uint32_t shots, uint64_t* out) {
  return sample_impl(c, st, nullptr, shots, out);
}

aleph_hip_status aleph_hip_sample_with_uniforms(aleph_hip_ctx* c, aleph_hip_state* st, const double* u, uint32_t shots,
                                                uint64_t* out) {
  if (shots && !u) return ALEPH_HIP_ERR_INVALID_ARGUMENT;
  return sample_impl(c, st, u, shots, out);
}

aleph_hip_status aleph_hip_sample_with_uniforms_noalloc(aleph_hip_ctx* c, aleph_hip_state* st, const double* u,
                                                        uint32_t shots, uint64_t* out) {
  return guarded(c, [&] {
    check_state(c, st);
    need(shots == 0 || (u != nullptr && out != nullptr), "null argument");
    if (st->n > kNoAllocMaxQubits || shots > kNoAllocMaxShots) {
      Failure f{ALEPH_HIP_ERR_INVALID_ARGUMENT,
                "allocation-free sampling takes at most 32 qubits and 65536 shots per call"};
      f.kind = "argument";
      throw f;
    }
    need(c->noalloc_scratch != nullptr, "sampler scratch not set up");
    validate(c, st, nullptr, nullptr, -1);   // context scratch, allocated with the context
    if (shots == 0) return;
    // Every region is written before it is read (block sums, their prefix and
    // total by the first two kernels, the uniforms by the copy, every result
    // slot by the shot kernel), so nothing needs clearing between calls.
    unsigned char* base = static_cast<unsigned char*>(c->noalloc_scratch);
    double* sums = reinterpret_cast<double*>(base + kNoAllocSumsOff);
    double* misc = reinterpret_cast<double*>(base + kNoAllocMiscOff);
    double* du = reinterpret_cast<double*>(base + kNoAllocUniOff);
    uint64_t* dout = reinterpret_cast<uint64_t*>(base + kNoAllocOutOff);
    ALEPH_HIP_CHECK(hipMemcpyAsync(du, u, shots * sizeof(double), hipMemcpyHostToDevice, c->stream));
    if (st->precision == ALEPH_HIP_F64)
      launch_sample<double>(c->stream, st->d, st->n, sums, misc, du, shots, dout);
    else
      launch_sample<float>(c->stream, st->d, st->n, sums, misc, du, shots, dout);
    ALEPH_HIP_CHECK(hipMemcpyAsync(out, dout, shots * sizeof(uint64_t), hipMemcpyDeviceToHost, c->stream));
    ALEPH_HIP_CHECK(hipStreamSynchronize(c->stream));
  });
}

aleph_hip_status aleph_hip_expectation(aleph_hip_ctx* c, aleph_hip_state* st, double coeff, const uint32_t* qubits,
                                       const uint8_t* paulis, uint32_t nt, double* out) {
  return guarded(c, [&] {
    check_state(c, st);
    need(out != nullptr && (nt == 0 || (qubits && paulis)), "null argument");
    validate(c, st, nullptr, nullptr, -1);
    if (!std::isfinite(coeff))
      throw make_failure(ALEPH_HIP_ERR_INVALID_PAULI_STRING, "non-finite coefficient", "invalid Pauli string");
    check_distinct(qubits, nt, st, "expectation");
    uint64_t x = 0, zy = 0;
    int ny = 0;
    for (uint32_t i = 0; i < nt; i++) {
      need(paulis[i] <= 3, "pauli code must be 0..3");
      const uint64_t b = 1ull << qubits[i];
      if (paulis[i] == 1) x |= b;
      if (paulis[i] == 2) { x |= b; zy |= b; ny++; }
      if (paulis[i] == 3) zy |= b;
    }
    // One term through the grouped kernel: for x != 0 it sums over the pairs
    // {i, i ^ x}, so every amplitude is read once (a per-element read of
    // psi[i] and psi[i ^ x] reads the state twice when x holds high qubits).
    double* partials = c->d_scratch;
    double* res = c->d_scratch + kResOffset;
    const uint64_t zys[1] = {zy};
    const bool odd = x != 0 && (__builtin_popcountll(x & zy) & 1);
    const uint8_t use_im[1] = {(uint8_t)odd};
    if (st->precision == ALEPH_HIP_F64)
      launch_expect_group<double>(c->stream, st->d, st->n, x, zys, use_im, 1u, partials, res);
    else
      launch_expect_group<float>(c->stream, st->d, st->n, x, zys, use_im, 1u, partials, res);
    ALEPH_HIP_CHECK(hipMemcpyAsync(c->h_pinned, res, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    ALEPH_HIP_CHECK(hipStreamSynchronize(c->stream));
    const double acc = c->h_pinned[0];
    const double re = x == 0 ? acc : (odd ? 0.0 : 2.0 * acc), im = odd ? 2.0 * acc : 0.0;
    double v;
    switch (ny & 3) {  // Re(i^ny * (re + i im))
      case 0: v = re; break;
      case 1: v = -im; break;
      case 2: v = -re; break;
      default: v = im; break;
    }
    *out = coeff * v;
  });
}

aleph_hip_status aleph_hip_expectation_sum(aleph_hip_ctx* c, aleph_hip_state* st, uint32_t n_terms,
                                           const double* coeffs, const uint32_t* offsets, const uint32_t* qubits,
                                           const uint8_t* paulis, double* out) {
  return guarded(c, [&] {
    check_state(c, st);
    need(out != nullptr && (n_terms == 0 || (coeffs && offsets)), "null argument");
    validate(c, st, nullptr, nullptr, -1);
    struct Term {
      uint64_t x, zy;
      int ny;
      double coeff;
    };
    std::vector<Term> terms(n_terms);
    for (uint32_t t = 0; t < n_terms; t++) {
      need(offsets[t] <= offsets[t + 1], "term offsets must be non-decreasing");
      const uint32_t nt = offsets[t + 1] - offsets[t];
      need(nt == 0 || (qubits && paulis), "null argument");
      if (!std::isfinite(coeffs[t]))
        throw make_failure(ALEPH_HIP_ERR_INVALID_PAULI_STRING, "non-finite coefficient", "invalid Pauli string");
      check_distinct(qubits + offsets[t], nt, st, "expectation");
      Term& tm = terms[t];
      tm = Term{0, 0, 0, coeffs[t]};
      for (uint32_t i = offsets[t]; i < offsets[t + 1]; i++) {
        need(paulis[i] <= 3, "pauli code must be 0..3");
        const uint64_t b = 1ull << qubits[i];
        if (paulis[i] == 1) tm.x |= b;
        if (paulis[i] == 2) { tm.x |= b; tm.zy |= b; tm.ny++; }
        if (paulis[i] == 3) tm.zy |= b;
      }
    }
    // Terms are grouped by X mask; each group of up to kExpectGroup terms is
    // one read of the state (pairs {i, i ^ x} read once each), so a
    // diagonal Hamiltonian such as a MaxCut sum costs one sweep per 16 terms.
    double* partials = c->d_scratch;
    double* res = c->d_scratch + kResOffset;
    std::vector<uint32_t> order(n_terms);
    for (uint32_t t = 0; t < n_terms; t++) order[t] = t;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return terms[a].x < terms[b].x; });
    struct Launch {
      uint64_t x;
      std::vector<uint32_t> ids;
    };
    std::vector<Launch> launches;
    for (uint32_t k = 0; k < n_terms; k++) {
      const Term& tm = terms[order[k]];
      if (tm.coeff == 0.0) continue;
      if (launches.empty() || launches.back().x != tm.x || launches.back().ids.size() == kExpectGroup)
        launches.push_back(Launch{tm.x, {}});
      launches.back().ids.push_back(order[k]);
    }
    // Results of up to 256 launches are read back per host synchronisation.
    const size_t per_sync = 256;
    std::vector<double> host(per_sync * kExpectGroup);
    double total = 0.0;
    for (size_t l0 = 0; l0 < launches.size(); l0 += per_sync) {
      const size_t nl = std::min(per_sync, launches.size() - l0);
      for (size_t j = 0; j < nl; j++) {
        const Launch& L = launches[l0 + j];
        uint64_t zy[kExpectGroup];
        uint8_t use_im[kExpectGroup];
        for (size_t t = 0; t < L.ids.size(); t++) {
          const Term& tm = terms[L.ids[t]];
          zy[t] = tm.zy;
          use_im[t] = (uint8_t)(tm.x != 0 && (__builtin_popcountll(tm.x & tm.zy) & 1));
        }
        double* out_j = res + j * kExpectGroup;
        if (st->precision == ALEPH_HIP_F64)
          launch_expect_group<double>(c->stream, st->d, st->n, L.x, zy, use_im, (uint32_t)L.ids.size(), partials,
                                      out_j);
        else
          launch_expect_group<float>(c->stream, st->d, st->n, L.x, zy, use_im, (uint32_t)L.ids.size(), partials,
                                     out_j);
      }
      ALEPH_HIP_CHECK(hipMemcpyAsync(host.data(), res, nl * kExpectGroup * sizeof(double), hipMemcpyDeviceToHost,
                                     c->stream));
      ALEPH_HIP_CHECK(hipStreamSynchronize(c->stream));
      for (size_t j = 0; j < nl; j++) {
        const Launch& L = launches[l0 + j];
        for (size_t t = 0; t < L.ids.size(); t++) {
          const Term& tm = terms[L.ids[t]];
          // (re, im) of sum_i conj(psi[i ^ x]) (-1)^{|i & zy|} psi[i]: the pair
          // sums count each pair once, so they carry a factor 2.
          double re = 0.0, im = 0.0;
          const double acc = host[j * kExpectGroup + t];
          if (tm.x == 0) re = acc;
          else if (__builtin_popcountll(tm.x & tm.zy) & 1) im = 2.0 * acc;
          else re = 2.0 * acc;
          double v;
          switch (tm.ny & 3) {  // Re(i^ny * (re + i im))
            case 0: v = re; break;
            case 1: v = -im; break;
            case 2: v = -re; break;
            default: v = im; break;
          }
          total += tm.coeff * v;
        }
      }
    }
    *out = total;
  });
}

aleph_hip_status aleph_hip_linear_xeb(aleph_hip_ctx* c, aleph_hip_state* st, double* out) {
  return guarded(c, [&] {
    check_state(c, st);
    need(out != nullptr, "null output");
    double* partials = c->d_scratch;
    double* res = c->d_scratch + kResOffset;
    if (st->precision == ALEPH_HIP_F64) launch_sum_p2<double>(c->stream, st->d, st->n, partials, res);
    else launch_sum_p2<float>(c->stream, st->d, st->n, partials, res);
    ALEPH_HIP_CHECK(hipMemcpyAsync(c->h_pinned, res, 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    ALEPH_HIP_CHECK(hipStreamSynchronize(c->stream));
    if (c->h_pinned[2] != 0.0) throw make_failure(ALEPH_HIP_ERR_INVALID_STATE, "non-finite amplitude", "invalid state");
    *out = std::ldexp(c->h_pinned[0], (int)st->n) - 1.0;
  });
}

aleph_hip_status aleph_hip_dist_alloc_rank(aleph_hip_ctx* c, uint32_t m, uint32_t precision, uint32_t rank,
                                           aleph_hip_state** out) {
  const aleph_hip_status s = aleph_hip_state_alloc(c, m, precision, out);
  if (s != ALEPH_HIP_OK) return s;
  if (rank == 0) return aleph_hip_state_reset(c, *out);
  return guarded(c, [&] {
    ALEPH_HIP_CHECK(hipMemsetAsync((*out)->d, 0, (precision == ALEPH_HIP_F64 ? 16ull : 8ull) << m, c->stream));
    ALEPH_HIP_CHECK(hipStreamSynchronize(c->stream));
  });
}

aleph_hip_status aleph_hip_gate_matrix(const aleph_hip_gate* g, uint32_t num_qubits, double* out, uint32_t* dim) {
  if (!g || !out || !dim) return ALEPH_HIP_ERR_INVALID_ARGUMENT;
  try {
    const DenseGate d = validate_gate(*g, num_qubits);
    *dim = d.dim;
    for (size_t e = 0; e < d.m.size(); e++) {
      out[2 * e] = d.m[e].real();
      out[2 * e + 1] = d.m[e].imag();
    }
    return ALEPH_HIP_OK;
  } catch (const Failure& f) {
    return f.code;
  }
}

aleph_hip_status aleph_hip_dist_exchange(aleph_hip_ctx* c, aleph_hip_state* const* ranks, uint32_t n_ranks,
                                         const uint32_t* global_bits, uint32_t k) {
  return guarded(c, [&] {
    need(c != nullptr && ranks != nullptr && global_bits != nullptr, "null argument");
    need(n_ranks >= 2 && (n_ranks & (n_ranks - 1)) == 0, "dist: rank count must be a power of two >= 2");
    const uint32_t g = (uint32_t)__builtin_ctz(n_ranks);
    for (uint32_t r = 0; r < n_ranks; r++) check_state(c, ranks[r]);
    const uint32_t m = ranks[0]->n, prec = ranks[0]->precision;
    for (uint32_t r = 1; r < n_ranks; r++)
      need(ranks[r]->n == m && ranks[r]->precision == prec && ranks[r] != ranks[0], "dist: rank slices differ");
    for (uint32_t a = 0; a < n_ranks; a++)
      for (uint32_t b = a + 1; b < n_ranks; b++) need(ranks[a] != ranks[b], "dist: a rank slice appears twice");
    need(k >= 1 && k <= m && k <= g && k <= kMaxExchangeBits, "dist: bad exchange bits");
    ExchangeBits bits{};
    uint32_t seen = 0;
    for (uint32_t j = 0; j < k; j++) {
      need(global_bits[j] >= m && global_bits[j] < m + g, "dist: bad exchange bits");
      const uint32_t rb = global_bits[j] - m;
      need(!((seen >> rb) & 1u), "dist: bad exchange bits");
      seen |= 1u << rb;
      bits.rank_bit[j] = rb;
    }
    std::vector<void*> ptrs(n_ranks);
    for (uint32_t r = 0; r < n_ranks; r++) ptrs[r] = ranks[r]->d;
    DevBuf dptr(n_ranks * sizeof(void*));
    ALEPH_HIP_CHECK(hipMemcpyAsync(dptr.p, ptrs.data(), n_ranks * sizeof(void*), hipMemcpyHostToDevice, c->stream));
    if (prec == ALEPH_HIP_F64) launch_exchange<double>(c->stream, (void* const*)dptr.p, n_ranks, m, k, bits);
    else launch_exchange<float>(c->stream, (void* const*)dptr.p, n_ranks, m, k, bits);
    ALEPH_HIP_CHECK(hipStreamSynchronize(c->stream));
  });
}

aleph_hip_status aleph_hip_run_noisy(aleph_hip_ctx* c, uint32_t n, const uint32_t* words, size_t n_words,
                                     const double* reals, size_t n_reals, const aleph_hip_channel* channels,
                                     uint32_t n_channels, const double* readout, uint64_t readout_mask,
                                     uint32_t shots, uint64_t seed, uint64_t shot_offset, size_t budget_bytes,
                                     uint64_t* outcomes) {
  return guarded(c, [&] {
    need(c != nullptr, "null context");
    need(shots == 0 || outcomes != nullptr, "null outcomes");
    need(n_words == 0 || words != nullptr, "null words");
    need(n_reals == 0 || reals != nullptr, "null reals");
    if (n < 1 || n > kNoiseMaxQubits) {
      Failure f = make_failure(ALEPH_HIP_ERR_TOO_MANY_QUBITS, "run_noisy", "noise trajectories need 1..24 qubits");
      f.expected = kNoiseMaxQubits;
      f.got = n;
      throw f;
    }
    const std::vector<ProgInstr> prog = decode_program(words, n_words, reals, n_reals, n);
    const TrajHost h = build_trajectories(prog, n, channels, n_channels, readout, readout_mask);
    if (shots == 0) return;
    uint32_t workgroups = 0;
    size_t slice_bytes = 0;
    trajectory_geometry(n, shots, budget_bytes ? budget_bytes : (size_t{4} << 30), &workgroups, &slice_bytes);
    const size_t op_bytes = h.ops.size() * sizeof(TrajOp), real_bytes = h.reals.size() * sizeof(double);
    const size_t mask_bytes = h.masks.size() * sizeof(uint64_t), index_bytes = h.index.size() * sizeof(uint32_t);
    const size_t ro_bytes = std::max<size_t>(h.readout.size(), 1) * sizeof(double);
    // One device block: program records, payloads, readout table, outcomes.
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t o_ops = 0, o_reals = al(op_bytes), o_masks = o_reals + al(real_bytes),
                 o_index = o_masks + al(mask_bytes), o_ro = o_index + al(index_bytes), o_out = o_ro + al(ro_bytes),
                 total = o_out + al((size_t)shots * sizeof(uint64_t));
    DevBuf blk(total);
    DevBuf slices(slice_bytes);
    unsigned char* base = (unsigned char*)blk.p;
    if (op_bytes)
      ALEPH_HIP_CHECK(hipMemcpyAsync(base + o_ops, h.ops.data(), op_bytes, hipMemcpyHostToDevice, c->stream));
    ALEPH_HIP_CHECK(hipMemcpyAsync(base + o_reals, h.reals.data(), real_bytes, hipMemcpyHostToDevice, c->stream));
    ALEPH_HIP_CHECK(hipMemcpyAsync(base + o_masks, h.masks.data(), mask_bytes, hipMemcpyHostToDevice, c->stream));
    ALEPH_HIP_CHECK(hipMemcpyAsync(base + o_index, h.index.data(), index_bytes, hipMemcpyHostToDevice, c->stream));
    if (!h.readout.empty())
      ALEPH_HIP_CHECK(hipMemcpyAsync(base + o_ro, h.readout.data(), h.readout.size() * sizeof(double),
                                     hipMemcpyHostToDevice, c->stream));
    TrajProgram tp;
    tp.ops = (const TrajOp*)(base + o_ops);
    tp.n_ops = (uint32_t)h.ops.size();
    tp.n = n;
    tp.reals = (const double*)(base + o_reals);
    tp.masks = (const uint64_t*)(base + o_masks);
    tp.index = (const uint32_t*)(base + o_index);
    tp.readout = (const double*)(base + o_ro);
    tp.readout_mask = h.readout_mask;
    tp.seed = seed;
    tp.shot0 = shot_offset;
    tp.shots = shots;
    uint64_t* d_out = (uint64_t*)(base + o_out);
    launch_trajectories(c->stream, tp, workgroups, slices.p, d_out);
    ALEPH_HIP_CHECK(hipMemcpyAsync(outcomes, d_out, (size_t)shots * sizeof(uint64_t), hipMemcpyDeviceToHost, c->stream));
    ALEPH_HIP_CHECK(hipStreamSynchronize(c->stream));
  });
}

aleph_hip_status aleph_hip_probabilities(aleph_hip_ctx* c, aleph_hip_state* st, const uint32_t* qubits, uint32_t nq,
                                         double* out) {
  return guarded(c, [&] {
    check_state(c, st);
    need(out != nullptr && (nq == 0 || qubits), "null argument");
    validate(c, st, nullptr, nullptr, -1);
    check_distinct(qubits, nq, st, "probabilities");
    if (nq == 0) {
      out[0] = 1.0;
      return;
    }
    need(nq <= 30, "at most 30 marginal qubits");
    const size_t bins = 1ull << nq;
    DevBuf d(bins * sizeof(double));
    if (st->precision == ALEPH_HIP_F64) launch_marginal<double>(c->stream, st->d, st->n, nq, qubits, (double*)d.p);
    else launch_marginal<float>(c->stream, st->d, st->n, nq, qubits, (double*)d.p);
    ALEPH_HIP_CHECK(hipMemcpyAsync(out, d.p, bins * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    ALEPH_HIP_CHECK(hipStreamSynchronize(c->stream));
  });
}

}  // extern "C"

// ---- batches ----

namespace {

BatchConfig batch_config(const aleph_hip_batch_opts* o) {
  BatchConfig c;
  if (!o) return c;
  need(o->mode == ALEPH_HIP_BATCH_PERSISTENT || o->mode == ALEPH_HIP_BATCH_ROUNDS, "unknown batch mode");
  c.mode = o->mode;
  c.local_qubits = o->local_qubits;
  c.chunks_per_item = o->chunks_per_item;
  c.groups_per_cu = o->groups_per_cu;
  c.max_ops = o->max_ops_per_pass;
  c.spin_limit = o->spin_limit;
  c.nontemporal = (o->flags & ALEPH_HIP_BATCH_NONTEMPORAL) != 0;
  c.unreachable_deps = (o->flags & ALEPH_HIP_BATCH_UNREACHABLE_DEPS) != 0;
  c.solo_qubits = o->solo_qubits;
  return c;
}

void check_batch_sizes(const uint32_t* nq, uint32_t n_inst) {
  need(n_inst > 0 && nq != nullptr, "empty batch");
  for (uint32_t i = 0; i < n_inst; i++) {
    if (nq[i] < 5 || nq[i] > 32) {
      Failure f = make_failure(ALEPH_HIP_ERR_TOO_MANY_QUBITS, "batch", "batch instances need 5..32 qubits");
      f.got = nq[i];
      f.expected = 32;
      throw f;
    }
  }
}

void finish_batch(aleph_hip_ctx* c, std::unique_ptr<Batch>& b, aleph_hip_batch** out) {
  b->plan();
  b->upload(c->stream, c->device);
  *out = new aleph_hip_batch{std::move(b)};
}

template <typename F>
void time_on_stream(hipStream_t s, double* ms, F&& body) {
  hipEvent_t e0, e1;
  ALEPH_HIP_CHECK(hipEventCreate(&e0));
  ALEPH_HIP_CHECK(hipEventCreate(&e1));
  ALEPH_HIP_CHECK(hipEventRecord(e0, s));
  body();
  ALEPH_HIP_CHECK(hipEventRecord(e1, s));
  ALEPH_HIP_CHECK(hipEventSynchronize(e1));
  float f = 0.f;
  ALEPH_HIP_CHECK(hipEventElapsedTime(&f, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *ms = f;
}

void run_checked(aleph_hip_ctx* c, Batch& b, double* ms) {
  if (ms) time_on_stream(c->stream, ms, [&] { b.run(c->stream); });
  else b.run(c->stream);
  if (b.failed(c->stream))
    throw make_failure(ALEPH_HIP_ERR_DEVICE_FAILURE, "batch",
                       "persistent batch: a dependency wait exceeded its spin limit; results are invalid");
}

}  // namespace

extern "C" {

aleph_hip_status aleph_hip_batch_create(aleph_hip_ctx* c, uint32_t n_inst, const uint32_t* nq, uint32_t precision,
                                        const uint32_t* words, const uint64_t* word_offsets, const double* reals,
                                        const uint64_t* real_offsets, const aleph_hip_batch_opts* opts,
                                        aleph_hip_batch** out) {
  return guarded(c, [&] {
    need(out != nullptr && word_offsets != nullptr && real_offsets != nullptr, "null argument");
    need(precision == ALEPH_HIP_F64 || precision == ALEPH_HIP_F32, "precision must be 64 or 32");
    check_batch_sizes(nq, n_inst);
    auto b = std::make_unique<Batch>();
    b->precision = precision;
    b->cfg = batch_config(opts);
    b->nq.assign(nq, nq + n_inst);
    b->progs.resize(n_inst);
    for (uint32_t i = 0; i < n_inst; i++) {
      need(word_offsets[i] <= word_offsets[i + 1] && real_offsets[i] <= real_offsets[i + 1], "bad offsets");
      b->progs[i] = decode_program(words + word_offsets[i], word_offsets[i + 1] - word_offsets[i],
                                   reals + real_offsets[i], real_offsets[i + 1] - real_offsets[i], nq[i]);
    }
    finish_batch(c, b, out);
  });
}

aleph_hip_status aleph_hip_batch_create_qaoa(aleph_hip_ctx* c, uint32_t n_inst, const uint32_t* nq,
                                             const uint64_t* edge_offsets, const uint32_t* edges, uint32_t p,
                                             const double* gammas, const double* betas, uint32_t precision,
                                             const aleph_hip_batch_opts* opts, aleph_hip_batch** out) {
  return guarded(c, [&] {
    need(out != nullptr && edge_offsets != nullptr && edges != nullptr, "null argument");
    need(p == 0 || (gammas != nullptr && betas != nullptr), "null angles");
    need(precision == ALEPH_HIP_F64 || precision == ALEPH_HIP_F32, "precision must be 64 or 32");
    check_batch_sizes(nq, n_inst);
    auto b = std::make_unique<Batch>();
    b->precision = precision;
    b->cfg = batch_config(opts);
    b->nq.assign(nq, nq + n_inst);
    b->qaoa.resize(n_inst);
    for (uint32_t i = 0; i < n_inst; i++) {
      need(edge_offsets[i] <= edge_offsets[i + 1], "bad edge offsets");
      QaoaSpec& q = b->qaoa[i];
      q.edges.assign(edges + 2 * edge_offsets[i], edges + 2 * edge_offsets[i + 1]);
      (void)cut_desc(nq[i], q.edges.data(), (uint32_t)(q.edges.size() / 2));  // validates the edges
      q.gammas.assign(gammas + (size_t)i * p, gammas + (size_t)(i + 1) * p);
      q.betas.assign(betas + (size_t)i * p, betas + (size_t)(i + 1) * p);
      for (uint32_t l = 0; l < p; l++)
        if (!std::isfinite(q.gammas[l]) || !std::isfinite(q.betas[l]))
          throw make_failure(ALEPH_HIP_ERR_NON_FINITE_PARAM, "qaoa", "non-finite angle");
    }
    finish_batch(c, b, out);
  });
}

aleph_hip_status aleph_hip_batch_set_qaoa_angles(aleph_hip_ctx* c, aleph_hip_batch* h, const double* gammas,
                                                 const double* betas) {
  return guarded(c, [&] {
    need(h != nullptr && h->b && !h->b->qaoa.empty(), "not a QAOA batch");
    Batch& b = *h->b;
    for (uint32_t i = 0; i < b.size(); i++) {
      QaoaSpec& q = b.qaoa[i];
      const size_t p = q.gammas.size();
      for (size_t l = 0; l < p; l++) {
        if (!std::isfinite(gammas[i * p + l]) || !std::isfinite(betas[i * p + l]))
          throw make_failure(ALEPH_HIP_ERR_NON_FINITE_PARAM, "qaoa", "non-finite angle");
        q.gammas[l] = gammas[i * p + l];
        q.betas[l] = betas[i * p + l];
      }
    }
    // Angles can change the plan's structure (a zero angle drops its gate),
    // so the captured graph is re-validated against the new plan on the next run.
    b.plan();
    b.upload(c->stream, c->device);
  });
}

aleph_hip_status aleph_hip_batch_run(aleph_hip_ctx* c, aleph_hip_batch* h) {
  return guarded(c, [&] {
    need(h != nullptr && h->b, "null batch");
    run_checked(c, *h->b, nullptr);
  });
}

aleph_hip_status aleph_hip_batch_run_timed(aleph_hip_ctx* c, aleph_hip_batch* h, double* ms) {
  return guarded(c, [&] {
    need(h != nullptr && h->b && ms != nullptr, "null argument");
    run_checked(c, *h->b, ms);
  });
}

aleph_hip_status aleph_hip_batch_expect_maxcut(aleph_hip_ctx* c, aleph_hip_batch* h, const uint64_t* edge_offsets,
                                               const uint32_t* edges, double* out) {
  return guarded(c, [&] {
    need(h != nullptr && h->b && edge_offsets != nullptr && edges != nullptr && out != nullptr, "null argument");
    Batch& b = *h->b;
    const uint32_t N = b.size();
    std::vector<CutDesc> descs(N);
    uint32_t blocks = 0;
    for (uint32_t i = 0; i < N; i++) {
      need(edge_offsets[i] <= edge_offsets[i + 1], "bad edge offsets");
      descs[i] = cut_desc(b.nq[i], edges + 2 * edge_offsets[i], (uint32_t)(edge_offsets[i + 1] - edge_offsets[i]));
      descs[i].state_off = b.offs[i];
      descs[i].block_begin = blocks;
      blocks += b.nq[i] > kBatchCutBlockBits ? 1u << (b.nq[i] - kBatchCutBlockBits) : 1u;
    }
    DevBuf d_desc(N * sizeof(CutDesc));
    DevBuf d_part((size_t)blocks * sizeof(double));
    DevBuf d_out(N * sizeof(double));
    ALEPH_HIP_CHECK(hipMemcpyAsync(d_desc.p, descs.data(), N * sizeof(CutDesc), hipMemcpyHostToDevice, c->stream));
    if (b.precision == ALEPH_HIP_F64)
      launch_batch_cut<double>(c->stream, b.pool, (const CutDesc*)d_desc.p, N, blocks, (double*)d_part.p,
                               (double*)d_out.p);
    else
      launch_batch_cut<float>(c->stream, b.pool, (const CutDesc*)d_desc.p, N, blocks, (double*)d_part.p,
                              (double*)d_out.p);
    ALEPH_HIP_CHECK(hipMemcpyAsync(out, d_out.p, N * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    ALEPH_HIP_CHECK(hipStreamSynchronize(c->stream));
  });
}

aleph_hip_status aleph_hip_batch_read_state(aleph_hip_ctx* c, aleph_hip_batch* h, uint32_t i, double* out) {
  return guarded(c, [&] {
    need(h != nullptr && h->b && out != nullptr, "null argument");
    Batch& b = *h->b;
    need(i < b.size(), "instance out of range");
    const uint64_t amps = 1ull << b.nq[i];
    if (b.precision == ALEPH_HIP_F64) {
      ALEPH_HIP_CHECK(hipMemcpyAsync(out, (const char*)b.pool + b.offs[i] * 16, amps * 16, hipMemcpyDeviceToHost,
                                     c->stream));
      ALEPH_HIP_CHECK(hipStreamSynchronize(c->stream));
    } else {
      std::vector<float> f(2 * amps);
      ALEPH_HIP_CHECK(hipMemcpyAsync(f.data(), (const char*)b.pool + b.offs[i] * 8, amps * 8, hipMemcpyDeviceToHost,
                                     c->stream));
      ALEPH_HIP_CHECK(hipStreamSynchronize(c->stream));
      for (size_t j = 0; j < 2 * amps; j++) out[j] = f[j];
    }
  });
}

aleph_hip_status aleph_hip_batch_stats_get(const aleph_hip_batch* h, aleph_hip_batch_stats* out) {
  if (!h || !h->b || !out) return ALEPH_HIP_ERR_INVALID_ARGUMENT;
  const Batch& b = *h->b;
  aleph_hip_batch_stats s{};
  s.plan_ms = b.plan_ms;
  s.sweep_bytes = b.sweep_bytes;
  s.pool_bytes = b.pool_bytes;
  s.n_instances = b.size();
  s.n_passes = b.total_passes;
  s.n_items = (uint32_t)b.items.size();
  s.n_rounds = b.round_begin.empty() ? 0 : (uint32_t)b.round_begin.size() - 1;
  s.grid = b.grid;
  s.threads = b.launch.threads;
  s.lds_bytes = (uint32_t)b.launch.lds_bytes;
  s.mode = b.cfg.mode;
  s.local_qubits = b.cfg.local_qubits;
  s.chunks_per_item = b.cfg.chunks_per_item;
  s.groups_per_cu = b.cfg.groups_per_cu;
  s.solo_instances = (uint32_t)b.solo.size();
  for (uint64_t mk : b.plus_mask) {
    s.prepared_instances += mk ? 1u : 0u;
    s.prepared_qubits += (uint32_t)__builtin_popcountll(mk);
  }
  *out = s;
  return ALEPH_HIP_OK;
}

aleph_hip_status aleph_hip_batch_autotune(aleph_hip_ctx* c, aleph_hip_batch* h, uint32_t reps,
                                          aleph_hip_batch_tune* results, uint32_t max_results, uint32_t* n_results) {
  return guarded(c, [&] {
    need(h != nullptr && h->b, "null batch");
    Batch& b = *h->b;
    reps = std::max(1u, reps);
    const BatchConfig base = b.cfg;
    std::vector<uint32_t> chunk_bits = {10u, 11u, 12u};
    const uint32_t cpis[] = {4, 16, 64};
    const uint32_t modes[] = {ALEPH_HIP_BATCH_PERSISTENT, ALEPH_HIP_BATCH_ROUNDS};
    std::vector<aleph_hip_batch_tune> all;
    aleph_hip_batch_tune best{};
    best.ms = 1e300;
    for (uint32_t m : chunk_bits) {
      for (uint32_t cpi : cpis) {
        for (uint32_t mode : modes) {
          b.cfg = base;
          b.cfg.local_qubits = m;
          b.cfg.chunks_per_item = cpi;
          b.cfg.mode = mode;
          b.drop_graph();
          b.plan();
          b.upload(c->stream, c->device);
          run_checked(c, b, nullptr);  // warm-up, also captures the graph
          double t = 1e300;
          for (uint32_t r = 0; r < reps; r++) {
            double ms = 0.0;
            run_checked(c, b, &ms);
            t = std::min(t, ms);
          }
          aleph_hip_batch_tune rec{mode, m, cpi, b.cfg.groups_per_cu, t};
          all.push_back(rec);
          if (t < best.ms) best = rec;
        }
      }
    }
    b.cfg = base;
    b.cfg.mode = best.mode;
    b.cfg.local_qubits = best.local_qubits;
    b.cfg.chunks_per_item = best.chunks_per_item;
    b.drop_graph();
    b.plan();
    b.upload(c->stream, c->device);
    if (n_results) *n_results = (uint32_t)all.size();
    if (results)
      for (uint32_t i = 0; i < std::min<uint32_t>(max_results, (uint32_t)all.size()); i++) results[i] = all[i];
  });
}

void aleph_hip_batch_free(aleph_hip_batch* h) { delete h; }

aleph_hip_status aleph_hip_batch_build_host_qaoa(uint32_t n_inst, const uint32_t* nq, const uint64_t* edge_offsets,
                                                 const uint32_t* edges, uint32_t p, const double* gammas,
                                                 const double* betas, uint32_t precision,
                                                 const aleph_hip_batch_opts* opts, aleph_hip_batch** out,
                                                 aleph_hip_error* err) {
  aleph_hip_ctx tmp{};
  const aleph_hip_status st = guarded(nullptr, [&] {
    need(out != nullptr && edge_offsets != nullptr && edges != nullptr, "null argument");
    need(p == 0 || (gammas != nullptr && betas != nullptr), "null angles");
    need(precision == ALEPH_HIP_F64 || precision == ALEPH_HIP_F32, "precision must be 64 or 32");
    check_batch_sizes(nq, n_inst);
    auto b = std::make_unique<Batch>();
    b->precision = precision;
    b->cfg = batch_config(opts);
    b->nq.assign(nq, nq + n_inst);
    b->qaoa.resize(n_inst);
    for (uint32_t i = 0; i < n_inst; i++) {
      QaoaSpec& q = b->qaoa[i];
      q.edges.assign(edges + 2 * edge_offsets[i], edges + 2 * edge_offsets[i + 1]);
      (void)cut_desc(nq[i], q.edges.data(), (uint32_t)(q.edges.size() / 2));
      q.gammas.assign(gammas + (size_t)i * p, gammas + (size_t)(i + 1) * p);
      q.betas.assign(betas + (size_t)i * p, betas + (size_t)(i + 1) * p);
    }
    try {
      b->plan();
    } catch (const Failure& f) {
      set_error(&tmp, f);
      throw;
    }
    *out = new aleph_hip_batch{std::move(b)};
  });
  if (err) *err = tmp.err;
  return st;
}

aleph_hip_status aleph_hip_batch_view_get(const aleph_hip_batch* h, aleph_hip_batch_view* out) {
  if (!h || !h->b || !out) return ALEPH_HIP_ERR_INVALID_ARGUMENT;
  const Batch& b = *h->b;
  aleph_hip_batch_view v{};
  v.items = b.items.data();
  v.n_items = (uint32_t)b.items.size();
  v.item_bytes = sizeof(BatchItem);
  v.passes = b.passes.data();
  v.n_passes = (uint32_t)b.passes.size();
  v.pass_bytes = sizeof(PassDev);
  v.stages = b.stages.data();
  v.n_stages = (uint32_t)b.stages.size();
  v.stage_bytes = sizeof(StageDev);
  v.ops = b.ops.data();
  v.n_ops = (uint32_t)b.ops.size();
  v.op_bytes = sizeof(DevOp);
  v.terms = b.terms.data();
  v.n_terms = (uint32_t)b.terms.size();
  v.term_bytes = sizeof(PhaseTerm);
  v.offsets = b.offs.data();
  v.n_instances = b.size();
  v.n_rounds = b.round_begin.empty() ? 0 : (uint32_t)b.round_begin.size() - 1;
  v.round_begin = b.round_begin.data();
  v.plus_masks = b.plus_mask.data();
  v.threads = b.launch.threads;
  v.lds_bytes = (uint32_t)b.launch.lds_bytes;
  v.total_amps = b.total_amps;
  *out = v;
  return ALEPH_HIP_OK;
}

// ---- cost model ----

aleph_hip_status aleph_hip_model_create(aleph_hip_ctx* c, const float* params, uint64_t seed, aleph_hip_model** out) {
  return guarded(c, [&] {
    need(out != nullptr, "null output");
    std::vector<float> p(kMlpParams, 0.f);
    if (params) {
      std::copy(params, params + kMlpParams, p.begin());
    } else {
      // He initialisation from SplitMix64 + Box-Muller; biases zero.
      uint64_t s = seed;
      auto u01 = [&] {
        s += 0x9E3779B97F4A7C15ull;
        uint64_t z = s;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return ((double)(z >> 11) + 0.5) * (1.0 / 9007199254740992.0);
      };
      auto normal = [&] { return std::sqrt(-2.0 * std::log(u01())) * std::cos(6.283185307179586 * u01()); };
      const int H = kMlpHidden, I = kMlpIn;
      size_t o = 0;
      auto fill = [&](int fan_in, int count) {
        const double sd = std::sqrt(2.0 / fan_in);
        for (int k = 0; k < count; k++) p[o++] = (float)(sd * normal());
      };
      fill(I, I * H);
      o += H;
      fill(H, H * H);
      o += H;
      fill(H, H * H);
      o += H;
      fill(H, H);
      o += 1;
    }
    auto m = std::make_unique<aleph_hip_model>();
    ALEPH_HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&m->d), kMlpParams * sizeof(float)));
    ALEPH_HIP_CHECK(hipMemcpy(m->d, p.data(), kMlpParams * sizeof(float), hipMemcpyHostToDevice));
    *out = m.release();
  });
}

aleph_hip_status aleph_hip_model_get_params(aleph_hip_ctx* c, const aleph_hip_model* m, float* out) {
  return guarded(c, [&] {
    need(m != nullptr && out != nullptr, "null argument");
    ALEPH_HIP_CHECK(hipStreamSynchronize(c->stream));
    ALEPH_HIP_CHECK(hipMemcpy(out, m->d, kMlpParams * sizeof(float), hipMemcpyDeviceToHost));
  });
}

aleph_hip_status aleph_hip_model_set_params(aleph_hip_ctx* c, aleph_hip_model* m, const float* params) {
  return guarded(c, [&] {
    need(m != nullptr && params != nullptr, "null argument");
    ALEPH_HIP_CHECK(hipStreamSynchronize(c->stream));
    ALEPH_HIP_CHECK(hipMemcpy(m->d, params, kMlpParams * sizeof(float), hipMemcpyHostToDevice));
  });
}

aleph_hip_status aleph_hip_model_predict(aleph_hip_ctx* c, const aleph_hip_model* m, const float* X, uint32_t rows,
                                         float* out) {
  return guarded(c, [&] {
    need(m != nullptr && (rows == 0 || (X != nullptr && out != nullptr)), "null argument");
    if (rows == 0) return;
    DevBuf dx((size_t)rows * kMlpIn * sizeof(float));
    DevBuf dy((size_t)rows * sizeof(float));
    ALEPH_HIP_CHECK(hipMemcpyAsync(dx.p, X, (size_t)rows * kMlpIn * sizeof(float), hipMemcpyHostToDevice,
                                   c->stream));
    launch_mlp_forward(c->stream, m->d, (const float*)dx.p, rows, (float*)dy.p);
    ALEPH_HIP_CHECK(hipMemcpyAsync(out, dy.p, rows * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    ALEPH_HIP_CHECK(hipStreamSynchronize(c->stream));
  });
}

aleph_hip_status aleph_hip_model_gradient(aleph_hip_ctx* c, const aleph_hip_model* m, const float* X, const float* y,
                                          uint32_t rows, float* grads, float* loss) {
  return guarded(c, [&] {
    need(m != nullptr && grads != nullptr && loss != nullptr && (rows == 0 || (X && y)), "null argument");
    DevBuf dx(std::max<size_t>(rows, 1) * kMlpIn * sizeof(float));
    DevBuf dy(std::max<size_t>(rows, 1) * sizeof(float));
    DevBuf dg(kMlpParams * sizeof(float));
    DevBuf dl(sizeof(float));
    if (rows) {
      ALEPH_HIP_CHECK(hipMemcpyAsync(dx.p, X, (size_t)rows * kMlpIn * sizeof(float), hipMemcpyHostToDevice,
                                     c->stream));
      ALEPH_HIP_CHECK(hipMemcpyAsync(dy.p, y, rows * sizeof(float), hipMemcpyHostToDevice, c->stream));
    }
    launch_mlp_grad(c->stream, m->d, (const float*)dx.p, (const float*)dy.p, rows, (float*)dg.p, (float*)dl.p);
    ALEPH_HIP_CHECK(hipMemcpyAsync(grads, dg.p, kMlpParams * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    ALEPH_HIP_CHECK(hipMemcpyAsync(loss, dl.p, sizeof(float), hipMemcpyDeviceToHost, c->stream));
    ALEPH_HIP_CHECK(hipStreamSynchronize(c->stream));
  });
}

aleph_hip_status aleph_hip_model_train(aleph_hip_ctx* c, aleph_hip_model* m, const float* X, const float* y,
                                       uint32_t rows, uint32_t epochs, uint32_t batch, float lr, uint64_t seed,
                                       float* loss) {
  return guarded(c, [&] {
    need(m != nullptr && loss != nullptr && rows > 0 && X != nullptr && y != nullptr, "null argument");
    need(batch > 0 && std::isfinite(lr) && lr > 0.f, "batch and learning rate must be positive");
    const size_t xb = (size_t)rows * kMlpIn * sizeof(float);
    DevBuf dx(xb), dy(rows * sizeof(float)), dg(kMlpParams * sizeof(float)), dl(sizeof(float));
    DevBuf dm(kMlpParams * sizeof(float)), dv(kMlpParams * sizeof(float));
    std::vector<float> zeros(kMlpParams, 0.f);
    ALEPH_HIP_CHECK(hipMemcpyAsync(dm.p, zeros.data(), kMlpParams * sizeof(float), hipMemcpyHostToDevice, c->stream));
    ALEPH_HIP_CHECK(hipMemcpyAsync(dv.p, zeros.data(), kMlpParams * sizeof(float), hipMemcpyHostToDevice, c->stream));
    std::vector<uint32_t> perm(rows);
    for (uint32_t i = 0; i < rows; i++) perm[i] = i;
    std::vector<float> xs((size_t)rows * kMlpIn), ys(rows);
    uint64_t s = seed;
    auto next = [&] {
      s += 0x9E3779B97F4A7C15ull;
      uint64_t z = s;
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
      return z ^ (z >> 31);
    };
    uint32_t t = 0;
    for (uint32_t e = 0; e < epochs; e++) {
      for (uint32_t i = rows - 1; i > 0; i--) std::swap(perm[i], perm[next() % (i + 1)]);
      for (uint32_t i = 0; i < rows; i++) {
        std::copy(X + (size_t)perm[i] * kMlpIn, X + (size_t)(perm[i] + 1) * kMlpIn, xs.begin() + (size_t)i * kMlpIn);
        ys[i] = y[perm[i]];
      }
      ALEPH_HIP_CHECK(hipMemcpyAsync(dx.p, xs.data(), xb, hipMemcpyHostToDevice, c->stream));
      ALEPH_HIP_CHECK(hipMemcpyAsync(dy.p, ys.data(), rows * sizeof(float), hipMemcpyHostToDevice, c->stream));
      for (uint32_t b0 = 0; b0 < rows; b0 += batch) {
        const uint32_t bs = std::min(batch, rows - b0);
        launch_mlp_grad(c->stream, m->d, (const float*)dx.p + (size_t)b0 * kMlpIn, (const float*)dy.p + b0, bs,
                        (float*)dg.p, (float*)dl.p);
        launch_mlp_adam(c->stream, m->d, (const float*)dg.p, (float*)dm.p, (float*)dv.p, lr, 0.9f, 0.999f, 1e-8f,
                        ++t);
      }
      ALEPH_HIP_CHECK(hipStreamSynchronize(c->stream));  // host buffers are rewritten next epoch
    }
    ALEPH_HIP_CHECK(hipMemcpyAsync(dx.p, X, xb, hipMemcpyHostToDevice, c->stream));
    ALEPH_HIP_CHECK(hipMemcpyAsync(dy.p, y, rows * sizeof(float), hipMemcpyHostToDevice, c->stream));
    launch_mlp_grad(c->stream, m->d, (const float*)dx.p, (const float*)dy.p, rows, (float*)dg.p, (float*)dl.p);
    ALEPH_HIP_CHECK(hipMemcpyAsync(loss, dl.p, sizeof(float), hipMemcpyDeviceToHost, c->stream));
    ALEPH_HIP_CHECK(hipStreamSynchronize(c->stream));
  });
}

void aleph_hip_model_free(aleph_hip_model* m) { delete m; }

aleph_hip_status aleph_hip_ctx_set_cost_model(aleph_hip_ctx* c, const aleph_hip_model* m) {
  return guarded(c, [&] {
    ALEPH_HIP_CHECK(hipStreamSynchronize(c->stream));
    if (!m) {
      if (c->d_model) ALEPH_HIP_CHECK(hipFree(c->d_model));
      c->d_model = nullptr;
      return;
    }
    if (!c->d_model) ALEPH_HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&c->d_model), kMlpParams * sizeof(float)));
    ALEPH_HIP_CHECK(hipMemcpy(c->d_model, m->d, kMlpParams * sizeof(float), hipMemcpyDeviceToDevice));
  });
}

aleph_hip_status aleph_hip_plan_features(const aleph_hip_plan* plan, float* out, uint32_t* launch, uint32_t max_rows,
                                         uint32_t* rows) {
  if (!plan || !plan->plan || !rows) return ALEPH_HIP_ERR_INVALID_ARGUMENT;
  std::vector<float> f;
  std::vector<uint32_t> l;
  plan_features(*plan->plan, f, l);
  *rows = (uint32_t)l.size();
  const uint32_t n = std::min<uint32_t>(max_rows, (uint32_t)l.size());
  if (out) std::copy(f.begin(), f.begin() + (size_t)n * kPlanFeatures, out);
  if (launch) std::copy(l.begin(), l.begin() + n, launch);
  return ALEPH_HIP_OK;
}

aleph_hip_status aleph_hip_plan_jit_source(const aleph_hip_plan* plan, uint32_t pass, char* out, size_t cap,
                                           size_t* len) {
  if (!plan || !plan->plan || !len || pass >= plan->plan->passes.size()) return ALEPH_HIP_ERR_INVALID_ARGUMENT;
  try {
    const std::string s = jit_pass_source(*plan->plan, pass);
    *len = s.size();
    if (out && cap) {
      const size_t n = std::min(cap - 1, s.size());
      std::memcpy(out, s.data(), n);
      out[n] = 0;
    }
    return ALEPH_HIP_OK;
  } catch (const Failure& f) {
    return f.code;
  }
}

}  // extern "C"
