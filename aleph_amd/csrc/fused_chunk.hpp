// Device-side body of the fused LDS-staged pass, shared by the one-pass
// launcher (k_fused.hip) and the batched executors (k_batch.hip).
//
// One chunk of 2^m amplitudes (the index bits outside the pass's chunk
// qubits fixed by the chunk number) is read from HBM once; every op of the
// pass is applied while it sits in registers (2^G amplitudes per thread) or
// is re-distributed through LDS between stages; it is written back once.
#pragma once

#include <type_traits>

#include "device_types.hpp"
#include "plan.hpp"

namespace aleph {
namespace fused {


// XOR-fold swizzle of the chunk-local index into LDS slots: bits >= SW are
// folded into the low SW bits, so lanes whose thread bits map to high local
// bits still spread over distinct 128-byte bank windows.  It is linear over
// XOR: swz(a ^ b) == swz(a) ^ swz(b).
template <int SW>
__device__ __forceinline__ uint32_t swz(uint32_t li) {
  const uint32_t h = li >> SW;
  return li ^ ((h ^ (h >> SW) ^ (h >> (2 * SW)) ^ (h >> (3 * SW))) & ((1u << SW) - 1u));
}

// x with its sign bit XORed by (s & t), s in {0, 0x80000000}.
__device__ __forceinline__ double flip_sign(double x, uint32_t s, uint32_t t) {
  return __builtin_bit_cast(double, __builtin_bit_cast(uint64_t, x) ^ ((uint64_t)(s & t) << 32));
}
__device__ __forceinline__ float flip_sign(float x, uint32_t s, uint32_t t) {
  return __builtin_bit_cast(float, __builtin_bit_cast(uint32_t, x) ^ (s & t));
}

template <typename T>
__device__ __forceinline__ C2<T> cmac2(T ar, T ai, C2<T> x, T br, T bi, C2<T> y) {
  C2<T> o;
  o.x = ar * x.x - ai * x.y + br * y.x - bi * y.y;
  o.y = ar * x.y + ai * x.x + br * y.y + bi * y.x;
  return o;
}

// Uniform global offset of register slot r (bit g of r selects rg[g]).
template <int G>
__device__ __forceinline__ uint64_t roff(const uint64_t (&rg)[G], int r) {
  uint64_t o = 0;
#pragma unroll
  for (int g = 0; g < G; g++)
    if (r & (1 << g)) o |= rg[g];
  return o;
}

template <int G>
__device__ __forceinline__ uint32_t loff(const uint32_t (&rl)[G], int r) {
  uint32_t o = 0;
#pragma unroll
  for (int g = 0; g < G; g++)
    if (r & (1 << g)) o |= rl[g];
  return o;
}

// CTL: kCtlNone / kCtlSome when the plan fixes whether the op has controls
// (kernels generated per pass), kCtlAny to decide from op->cmask.
constexpr int kCtlNone = 0, kCtlSome = 1, kCtlAny = 2;

template <typename T, int G, int B, int CTL = kCtlAny>
__device__ __forceinline__ void op_1q(C2<T> (&a)[1 << G], const DevOp* __restrict__ op, uint64_t tgo,
                                      const uint64_t (&rg)[G]) {
  const T* mm = reinterpret_cast<const T*>(op->m);  // matrix stored in the plan's precision
  const T m00r = mm[0], m00i = mm[1], m01r = mm[2], m01i = mm[3];
  const T m10r = mm[4], m10i = mm[5], m11r = mm[6], m11i = mm[7];
  const uint64_t cm = op->cmask;
  if (CTL == kCtlNone || (CTL == kCtlAny && cm == 0)) {
#pragma unroll
    for (int r = 0; r < (1 << G); r++) {
      if (r & (1 << B)) continue;
      const int s = r | (1 << B);
      const C2<T> x = a[r], y = a[s];
      a[r] = cmac2(m00r, m00i, x, m01r, m01i, y);
      a[s] = cmac2(m10r, m10i, x, m11r, m11i, y);
    }
  } else {
    const uint64_t tc = tgo & cm;
#pragma unroll
    for (int r = 0; r < (1 << G); r++) {
      if (r & (1 << B)) continue;
      const int s = r | (1 << B);
      const C2<T> x = a[r], y = a[s];
      const bool on = (tc | (roff<G>(rg, r) & cm)) == cm;
      const C2<T> nx = cmac2(m00r, m00i, x, m01r, m01i, y);
      const C2<T> ny = cmac2(m10r, m10i, x, m11r, m11i, y);
      a[r] = on ? nx : x;
      a[s] = on ? ny : y;
    }
  }
}

// Insert a zero bit at position p of x (compile time).
constexpr int ins0(int x, int p) { return (x & ((1 << p) - 1)) | ((x >> p) << (p + 1)); }

// One quartet (register bits A > B cleared in slot Q): its four inputs are
// copied out, then each output row is formed from one matrix row read just
// before use.  For FP64 the empty asm with a memory clobber keeps the
// compiler from hoisting all sixteen matrix entries into registers at once.
template <typename T, int G, int A, int B, int Q, bool CTL>
__device__ __forceinline__ void quartets_2q(C2<T> (&a)[1 << G], const DevOp* __restrict__ op,
                                            const bool (&on)[1 << (G - 2)]) {
  if constexpr (Q < (1 << (G - 2))) {
    constexpr int r = ins0(ins0(Q, B), A);
    constexpr int idx[4] = {r, r | (1 << B), r | (1 << A), r | (1 << A) | (1 << B)};
    const C2<T> v0 = a[idx[0]], v1 = a[idx[1]], v2 = a[idx[2]], v3 = a[idx[3]];
    C2<T> out[4];
#pragma unroll
    for (int row = 0; row < 4; row++) {
      if constexpr (sizeof(T) == 8) asm volatile("" ::: "memory");
      const T* mrow = reinterpret_cast<const T*>(op->m) + 8 * row;
      const T m0r = mrow[0], m0i = mrow[1], m1r = mrow[2], m1i = mrow[3];
      const T m2r = mrow[4], m2i = mrow[5], m3r = mrow[6], m3i = mrow[7];
      T re = m0r * v0.x;
      T im = m0r * v0.y;
      re = fma(-m0i, v0.y, re);
      im = fma(m0i, v0.x, im);
      re = fma(m1r, v1.x, re);
      re = fma(-m1i, v1.y, re);
      im = fma(m1r, v1.y, im);
      im = fma(m1i, v1.x, im);
      re = fma(m2r, v2.x, re);
      re = fma(-m2i, v2.y, re);
      im = fma(m2r, v2.y, im);
      im = fma(m2i, v2.x, im);
      re = fma(m3r, v3.x, re);
      re = fma(-m3i, v3.y, re);
      im = fma(m3r, v3.y, im);
      im = fma(m3i, v3.x, im);
      out[row].x = re;
      out[row].y = im;
    }
    if constexpr (CTL) {
      const bool live = on[Q];
      a[idx[0]] = live ? out[0] : v0;
      a[idx[1]] = live ? out[1] : v1;
      a[idx[2]] = live ? out[2] : v2;
      a[idx[3]] = live ? out[3] : v3;
    } else {
#pragma unroll
      for (int k = 0; k < 4; k++) a[idx[k]] = out[k];
    }
    quartets_2q<T, G, A, B, Q + 1, CTL>(a, op, on);
  }
}

template <typename T, int G, int A, int B, int CTL = kCtlAny>  // A = matrix MSB register bit, B = LSB, A > B
__device__ __forceinline__ void op_2q(C2<T> (&a)[1 << G], const DevOp* __restrict__ op, uint64_t tgo,
                                      const uint64_t (&rg)[G]) {
  const uint64_t cm = op->cmask;
  bool on[1 << (G - 2)];
  if (CTL == kCtlNone || (CTL == kCtlAny && cm == 0)) {
    quartets_2q<T, G, A, B, 0, false>(a, op, on);
    return;
  }
  const uint64_t tc = tgo & cm;
#pragma unroll
  for (int q = 0; q < (1 << (G - 2)); q++) on[q] = (tc | (roff<G>(rg, ins0(ins0(q, B), A)) & cm)) == cm;
  quartets_2q<T, G, A, B, 0, true>(a, op, on);
}

// FP32 4x4 op on the matrix cores: v_mfma_f32_4x4x1f32 (16 blocks of
// 4 lanes).  In block b, lane 4b + n supplies A[n] and B[n] and receives
// column n of C += A B^T, i.e. C[i] += A(lane 4b + i) * B(own).  With each
// lane supplying row n of the op's matrix column k as A and its own
// quartet's element k as B, four K steps leave (M x)[i] in the lane's C[i]:
// every lane's quartet is its own column and nothing moves between lanes.
// Real and imaginary parts: yr = Mr xr - Mi xi, yi = Mi xr + Mr xi, eight
// K steps each, so 16 MFMAs per quartet replace 64 VALU FMAs, on the
// matrix pipe beside the vector ALU.  The lane's matrix row (8 floats of
// the plan's FP32 op matrix) is read with two 16-byte loads.
typedef float v4f32 __attribute__((ext_vector_type(4)));

// All quartets of the op advance together through the K steps, so the
// matrix pipe always has 2 * 2^(G-2) independent accumulation chains.
template <int G, int A, int B>
__device__ __forceinline__ void quartets_2q_mfma(C2<float> (&a)[1 << G], const float (&re)[4], const float (&im)[4],
                                                 const float (&nim)[4]) {
  constexpr int NQ = 1 << (G - 2);
  v4f32 cr[NQ], ci[NQ];
#pragma unroll
  for (int q = 0; q < NQ; q++) {
    cr[q] = v4f32{0.f, 0.f, 0.f, 0.f};
    ci[q] = v4f32{0.f, 0.f, 0.f, 0.f};
  }
#pragma unroll
  for (int k = 0; k < 4; k++) {
#pragma unroll
    for (int q = 0; q < NQ; q++) {
      const int r = ins0(ins0(q, B), A);
      const int slot = r | ((k & 1) << B) | ((k >> 1) << A);
      cr[q] = __builtin_amdgcn_mfma_f32_4x4x1f32(re[k], a[slot].x, cr[q], 0, 0, 0);
      ci[q] = __builtin_amdgcn_mfma_f32_4x4x1f32(im[k], a[slot].x, ci[q], 0, 0, 0);
    }
#pragma unroll
    for (int q = 0; q < NQ; q++) {
      const int r = ins0(ins0(q, B), A);
      const int slot = r | ((k & 1) << B) | ((k >> 1) << A);
      cr[q] = __builtin_amdgcn_mfma_f32_4x4x1f32(nim[k], a[slot].y, cr[q], 0, 0, 0);
      ci[q] = __builtin_amdgcn_mfma_f32_4x4x1f32(re[k], a[slot].y, ci[q], 0, 0, 0);
    }
  }
#pragma unroll
  for (int q = 0; q < NQ; q++) {
    const int r = ins0(ins0(q, B), A);
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int slot = r | ((i & 1) << B) | ((i >> 1) << A);
      a[slot].x = cr[q][i];
      a[slot].y = ci[q][i];
    }
  }
}

// A = matrix MSB register bit, B = LSB, A > B; uncontrolled ops only.
template <int G, int A, int B>
__device__ __forceinline__ void op_2q_mfma(C2<float> (&a)[1 << G], const DevOp* __restrict__ op) {
  const uint32_t n = threadIdx.x & 3u;          // this lane's matrix row within its 4-lane block
  const float* row = reinterpret_cast<const float*>(op->m) + 8 * n;
  const v4f32 lo = *reinterpret_cast<const v4f32*>(row);       // M[n][0], M[n][1] (re, im interleaved)
  const v4f32 hi = *reinterpret_cast<const v4f32*>(row + 4);   // M[n][2], M[n][3]
  const float re[4] = {lo[0], lo[2], hi[0], hi[2]};
  const float im[4] = {lo[1], lo[3], hi[1], hi[3]};
  const float nim[4] = {-lo[1], -lo[3], -hi[1], -hi[3]};
  quartets_2q_mfma<G, A, B>(a, re, im, nim);
}

template <typename T, int G>
__device__ __forceinline__ void op_diag(C2<T> (&a)[1 << G], const DevOp* __restrict__ op, uint64_t tgo,
                                        const uint64_t (&rg)[G]) {
  const uint64_t ma = op->ma, mb = op->mb, cm = op->cmask;
  const T* dm = reinterpret_cast<const T*>(op->m);
  const T d0r = dm[0], d0i = dm[1], d1r = dm[2], d1i = dm[3];
  const T d2r = dm[4], d2i = dm[5], d3r = dm[6], d3i = dm[7];
#pragma unroll
  for (int r = 0; r < (1 << G); r++) {
    const uint64_t idx = tgo | roff<G>(rg, r);
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

// RM >= 0: the register part of the mask fixed at compile time (generated
// pass kernels), so only the matching slots are touched; RM < 0: op->a.
template <typename T, int G, int RM = -1>
__device__ __forceinline__ void op_sign(C2<T> (&a)[1 << G], const DevOp* __restrict__ op, uint64_t tgo) {
  // Sign-bit XOR with (thread condition as a bit mask) & (uniform slot
  // condition): no divergent branch, one bitop per word.
  const uint64_t mo = op->cmask;
  const uint32_t rm = RM >= 0 ? (uint32_t)RM : op->a;
  const uint32_t tm = ((tgo & mo) == mo) ? ~0u : 0u;
#pragma unroll
  for (int r = 0; r < (1 << G); r++) {
    if (RM >= 0 && ((uint32_t)r & (uint32_t)RM) != (uint32_t)RM) continue;
    const uint32_t sr = (((uint32_t)r & rm) == rm) ? 0x80000000u : 0u;   // uniform
    a[r].x = flip_sign(a[r].x, sr, tm);
    a[r].y = flip_sign(a[r].y, sr, tm);
  }
}

// e^{i phi} for a phase op on amplitudes of type T.  The angle sums stay in
// FP64; an FP32 state takes the factor from the FP32 sincosf of the angle
// reduced to [-pi, pi] in FP64 (error ~1e-7, the state's own rounding),
// which costs a fraction of the FP64 sincos.
template <typename T>
__device__ __forceinline__ void phase_sincos(double phi, double* s, double* c) {
  if constexpr (std::is_same<T, float>::value) {
    constexpr double kTwoPi = 6.283185307179586476925286766559;
    const float r = (float)(phi - kTwoPi * rint(phi * (1.0 / kTwoPi)));
    float sf, cf;
    sincosf(r, &sf, &cf);
    *s = sf;
    *c = cf;
  } else {
    sincos(phi, s, c);
  }
}

// Sum of the angles of terms [b, e) that fire for this thread.  The loads
// are independent, so unrolling keeps several scalar loads in flight.
__device__ __forceinline__ double sum_terms(const PhaseTerm* __restrict__ terms, uint32_t b, uint32_t e,
                                            uint64_t tgo) {
  double s = 0.0;
#pragma unroll 4
  for (uint32_t t = b; t < e; t++) {
    const uint64_t mo = terms[t].mask_out;
    s += ((tgo & mo) == mo) ? terms[t].angle : 0.0;
  }
  return s;
}

template <typename T, int G>
__device__ __forceinline__ void op_phase(C2<T> (&a)[1 << G], const DevOp* __restrict__ op,
                                         const PhaseTerm* __restrict__ terms, uint64_t tgo) {
  // Terms on at most two register bits fold into 1 + G + G(G-1)/2
  // accumulators (PhaseSlots).  The host sorts them by accumulator and
  // stores the class boundaries (relative to term_off) in the op's matrix
  // slots; the rare terms on three or more register bits follow (op->b of
  // them) and are evaluated per amplitude.
  using L = PhaseSlots<G>;
  const uint32_t* bounds = reinterpret_cast<const uint32_t*>(op->m);
  double acc[L::kCount];
#pragma unroll
  for (int k = 0; k < L::kCount; k++)
    acc[k] = sum_terms(terms, op->term_off + bounds[k], op->term_off + bounds[k + 1], tgo);
  const uint32_t end = op->term_off + op->n_terms;
  const uint32_t gen_end = end + op->b;
  if (op->b == 0) {
    // No terms on three or more register bits: e^{i phi_r} is a product of
    // the class exponentials.  One sincos per class (1 + G + G(G-1)/2)
    // instead of one per amplitude; slot r's factor is built from slot
    // r without its top bit times e^{i(linear(top) + sum pair(g, top))}
    // over the remaining bits g, i.e. popcount(r) complex products.
    // Classes without terms (the common case: a controlled-phase ladder
    // fills 4 or 5 of the 11) contribute the factor 1: no sincos.
    double ec[L::kCount], es[L::kCount];
#pragma unroll
    for (int k = 0; k < L::kCount; k++) {
      if (bounds[k] == bounds[k + 1]) {
        ec[k] = 1.0;
        es[k] = 0.0;
      } else {
        phase_sincos<T>(acc[k], &es[k], &ec[k]);
      }
    }
    double fr[1 << G], fi[1 << G];
    fr[0] = ec[0];
    fi[0] = es[0];
#pragma unroll
    for (int r = 1; r < (1 << G); r++) {
      int top = 0;
#pragma unroll
      for (int g = 0; g < G; g++)
        if (r & (1 << g)) top = g;
      const int rest = r & ~(1 << top);
      double cr = ec[L::linear(top)], ci = es[L::linear(top)];
#pragma unroll
      for (int g = 0; g < G; g++) {
        if (!(rest & (1 << g))) continue;
        const double pr = ec[L::pair(g, top)], pi = es[L::pair(g, top)];
        const double t = cr * pr - ci * pi;
        ci = cr * pi + ci * pr;
        cr = t;
      }
      fr[r] = fr[rest] * cr - fi[rest] * ci;
      fi[r] = fr[rest] * ci + fi[rest] * cr;
    }
#pragma unroll
    for (int r = 0; r < (1 << G); r++) {
      const C2<T> x = a[r];
      a[r].x = x.x * (T)fr[r] - x.y * (T)fi[r];
      a[r].y = x.x * (T)fi[r] + x.y * (T)fr[r];
    }
    return;
  }
#pragma unroll
  for (int r = 0; r < (1 << G); r++) {
    double phi = acc[0];
#pragma unroll
    for (int g = 0; g < G; g++) {
      if (r & (1 << g)) phi += acc[L::linear(g)];
#pragma unroll
      for (int h = g + 1; h < G; h++)
        if ((r & (1 << g)) && (r & (1 << h))) phi += acc[L::pair(g, h)];
    }
    for (uint32_t t = end; t < gen_end; t++) {
      const PhaseTerm* pt = terms + t;
      const uint64_t mo = pt->mask_out;
      const uint32_t rm = pt->rmask;
      if ((tgo & mo) == mo && ((uint32_t)r & rm) == rm) phi += pt->angle;
    }
    double s, c;
    phase_sincos<T>(phi, &s, &c);
    const C2<T> x = a[r];
    a[r].x = x.x * (T)c - x.y * (T)s;
    a[r].y = x.x * (T)s + x.y * (T)c;
  }
}

// ---- phase ops in generated pass kernels ----
// The host orders each class's terms uniform-first (terms that test only
// bits outside the chunk, the same for every thread of a workgroup) and
// records the split (PhaseSlots::kCount + 1 class bounds, then kCount
// splits, in the op's matrix slots).  Generated kernels evaluate every term
// of the pass's phase ops for the chunk once per workgroup (one lane per
// term, into LDS), sum each class's uniform terms and take e^{i sum} (one
// lane per class, into LDS), and per thread multiply in only the unit
// factors of the thread-dependent terms: no per-thread sincos and no
// per-thread pass over the uniform terms.

// Contribution of the op's term `rel` (relative to term_off) for the chunk
// whose outside-chunk index bits are `base`: its angle if it fires.  Only
// the workgroup-uniform terms' values are used (phase_class_uniform).
__device__ __forceinline__ double phase_term_uniform(const DevOp* __restrict__ op,
                                                     const PhaseTerm* __restrict__ terms, uint64_t base,
                                                     uint32_t rel) {
  const PhaseTerm& pt = terms[op->term_off + rel];
  return ((base & pt.mask_out) == pt.mask_out) ? pt.angle : 0.0;
}

// Uniform factor of class k: e^{i sum} over the class's uniform terms,
// whose per-chunk angles `ua` (indexed relative to term_off) are in LDS.
template <int G>
__device__ __forceinline__ double2 phase_class_uniform(const DevOp* __restrict__ op, const double* ua, uint32_t k) {
  using L = PhaseSlots<G>;
  const uint32_t* bounds = reinterpret_cast<const uint32_t*>(op->m);
  const uint32_t* split = bounds + L::kCount + 1;
  double acc = 0.0;
  for (uint32_t t = bounds[k]; t < split[k]; t++) acc += ua[t];
  double sn, cs;
  sincos(acc, &sn, &cs);
  return make_double2(cs, sn);
}

// Factor of one class for this thread: the workgroup's uniform factor `u`
// times the unit factors of the class's COUNT thread-dependent terms, which
// start REL terms after the op's first term.
// The unit factors are selected independently and multiplied as a balanced
// tree, so the dependent chain is log2(COUNT) + 1 products long instead of
// COUNT: these passes run at two waves per SIMD and hide little latency.
template <int REL, int COUNT>
__device__ __forceinline__ double2 phase_factor(const DevOp* __restrict__ op, const PhaseTerm* __restrict__ terms,
                                                uint64_t tgo, double2 u) {
  if constexpr (COUNT == 0) {
    return u;
  } else {
    const PhaseTerm* tp = terms + op->term_off + REL;
    double fc[COUNT], fs[COUNT];
#pragma unroll
    for (int j = 0; j < COUNT; j++) {
      const uint64_t mo = tp[j].mask_out;
      const bool on = (tgo & mo) == mo;
      fc[j] = on ? tp[j].c : 1.0;
      fs[j] = on ? tp[j].s : 0.0;
    }
#pragma unroll
    for (int w = 1; w < COUNT; w *= 2) {
#pragma unroll
      for (int j = 0; j + w < COUNT; j += 2 * w) {
        const double r = fc[j] * fc[j + w] - fs[j] * fs[j + w];
        fs[j] = fc[j] * fs[j + w] + fs[j] * fc[j + w];
        fc[j] = r;
      }
    }
    return make_double2(u.x * fc[0] - u.y * fs[0], u.x * fs[0] + u.y * fc[0]);
  }
}

template <typename T>
__device__ __forceinline__ void phase_mul(C2<T>& a, double2 f) {
  const T fr = (T)f.x, fi = (T)f.y;
  const T x = a.x, y = a.y;
  a.x = x * fr - y * fi;
  a.y = x * fi + y * fr;
}

template <typename T, int G, int B>
__device__ __forceinline__ void dispatch_1q(C2<T> (&a)[1 << G], const DevOp* op, uint64_t tgo,
                                            const uint64_t (&rg)[G], uint32_t b) {
  if constexpr (B < G) {
    if (b == (uint32_t)B) op_1q<T, G, B>(a, op, tgo, rg);
    else dispatch_1q<T, G, B + 1>(a, op, tgo, rg, b);
  }
}

template <typename T, int G, int A, int B>
__device__ __forceinline__ void dispatch_2q(C2<T> (&a)[1 << G], const DevOp* op, uint64_t tgo,
                                            const uint64_t (&rg)[G], uint32_t ab) {
  if constexpr (A < G) {
    if constexpr (B < A) {
      if (ab == (uint32_t)(A * 8 + B)) op_2q<T, G, A, B>(a, op, tgo, rg);
      else dispatch_2q<T, G, A, B + 1>(a, op, tgo, rg, ab);
    } else {
      dispatch_2q<T, G, A + 1, 0>(a, op, tgo, rg, ab);
    }
  }
}

template <typename T, int G, int FEAT>
__device__ __forceinline__ void run_ops(C2<T> (&a)[1 << G], const DevOp* __restrict__ ops,
                                        const PhaseTerm* __restrict__ terms, uint32_t begin, uint32_t end,
                                        uint64_t tgo, const uint64_t (&rg)[G]) {
  for (uint32_t o = begin; o < end; o++) {
    const DevOp* op = ops + o;
    const uint32_t type = op->type;
    if (type == OP_1Q) {
      dispatch_1q<T, G, 0>(a, op, tgo, rg, op->a);
    } else if (type == OP_SIGN) {
      op_sign<T, G>(a, op, tgo);
    } else if (type == OP_DIAG) {
      op_diag<T, G>(a, op, tgo, rg);
    } else if (type == OP_2Q) {
      if constexpr ((FEAT & kFeat2Q) != 0) dispatch_2q<T, G, 1, 0>(a, op, tgo, rg, op->a * 8 + op->b);
    } else {
      if constexpr ((FEAT & kFeatPhase) != 0) op_phase<T, G>(a, op, terms, tgo);
    }
  }
}

typedef double v2f64 __attribute__((ext_vector_type(2)));
typedef float v2f32 __attribute__((ext_vector_type(2)));

template <bool NT, typename V>
__device__ __forceinline__ V gload(const V* p) {
  if constexpr (NT) return __builtin_nontemporal_load(p);
  else return *p;
}

template <bool NT, typename V>
__device__ __forceinline__ void gstore(V* p, V v) {
  if constexpr (NT) __builtin_nontemporal_store(v, p);
  else *p = v;
}

// Process chunk `chunk` of pass `pd` on `state`.  With GUARD, threads with
// tid >= 2^(m - G) hold no amplitudes but take part in every barrier, so a
// workgroup sized for the largest chunk can run smaller passes; without it
// the workgroup has exactly 2^(m - G) threads.
template <typename T, int G, int FEAT, bool NT, bool GUARD>
__device__ __forceinline__ void fused_chunk(C2<T>* __restrict__ state, const PassDev* __restrict__ pd,
                                            const StageDev* __restrict__ stages, const DevOp* __restrict__ ops,
                                            const PhaseTerm* __restrict__ terms, uint64_t chunk,
                                            C2<T>* __restrict__ lds) {
  constexpr int SW = sizeof(T) == 8 ? 3 : 4;
  constexpr int N = 1 << G;
  using V = typename std::conditional<sizeof(T) == 8, v2f64, v2f32>::type;
  const uint32_t m = pd->m;
  uint64_t base = chunk;
  for (uint32_t i = 0; i < m; i++) base = insert_zero(base, pd->local_pos[i]);
  const uint32_t tid = threadIdx.x;
  const bool active = !GUARD || tid < (1u << (m - G));
  const uint32_t ns = pd->n_stages;
  C2<T> a[N];
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
    uint64_t rg[G];
    uint32_t rl[G];
#pragma unroll
    for (int g = 0; g < G; g++) {
      rg[g] = sd->rg[g];
      rl[g] = sd->rl[g];
    }
    // Thread and register bits are disjoint, so the global index is the
    // thread's base plus a uniform register offset (sd->go), and (swz being
    // linear over XOR) the LDS slot is swz(thread part) ^ swz(register part),
    // the latter precomputed per stage (sd->ls): one vector op per access.
    C2<T>* const tp = state + tgo;
    const uint32_t ts = swz<SW>(tlo);
    if (active) {
      if (s == 0) {
#pragma unroll
        for (int r = 0; r < N; r++) {
          const V v = gload<NT>(reinterpret_cast<const V*>(tp + sd->go[r]));
          a[r].x = v.x;
          a[r].y = v.y;
        }
      } else {
#pragma unroll
        for (int r = 0; r < N; r++) a[r] = lds[ts ^ sd->ls[r]];
      }
      run_ops<T, G, FEAT>(a, ops, terms, sd->op_begin, sd->op_end, tgo, rg);
    }
    if (s + 1 == ns) {
      if (active) {
#pragma unroll
        for (int r = 0; r < N; r++) {
          V v;
          v.x = a[r].x;
          v.y = a[r].y;
          gstore<NT>(reinterpret_cast<V*>(tp + sd->go[r]), v);
        }
      }
    } else {
      if (s > 0) __syncthreads();
      if (active) {
#pragma unroll
        for (int r = 0; r < N; r++) lds[ts ^ sd->ls[r]] = a[r];
      }
      __syncthreads();
    }
  }
}

// Stage-by-stage access to one chunk, for pass kernels generated with the
// pass's op sequence written out (one straight-line body per pass).  The
// same address maps as fused_chunk: enter(s) selects stage s's thread and
// register bit maps; load/store move the chunk between HBM and registers;
// put/get move it between registers and LDS in the current stage's layout.
template <typename T, int G, bool NT>
struct ChunkIO {
  static constexpr int SW = sizeof(T) == 8 ? 3 : 4;
  static constexpr int N = 1 << G;
  using V = typename std::conditional<sizeof(T) == 8, v2f64, v2f32>::type;
  C2<T>* state;
  const StageDev* stages;
  C2<T>* lds;
  uint64_t base;
  uint64_t tgo;
  uint64_t rg[G];
  uint32_t rl[G];
  uint32_t ts;
  C2<T>* tp;
  const StageDev* cur;

  __device__ __forceinline__ ChunkIO(C2<T>* st, const PassDev* pd, const StageDev* sg, C2<T>* l, uint64_t chunk)
      : state(st), stages(sg), lds(l) {
    base = chunk;
    const uint32_t m = pd->m;
    for (uint32_t i = 0; i < m; i++) base = insert_zero(base, pd->local_pos[i]);
  }
  __device__ __forceinline__ void enter(uint32_t s) {
    const StageDev* sd = stages + s;
    cur = sd;
    const uint32_t tid = threadIdx.x;
    tgo = base;
    uint32_t tlo = 0;
    const uint32_t ntb = sd->n_thread_bits;
    for (uint32_t j = 0; j < ntb; j++) {
      if ((tid >> j) & 1u) {
        tgo |= sd->tg[j];
        tlo |= sd->tl[j];
      }
    }
#pragma unroll
    for (int g = 0; g < G; g++) {
      rg[g] = sd->rg[g];
      rl[g] = sd->rl[g];
    }
    tp = state + tgo;
    ts = swz<SW>(tlo);
  }
  __device__ __forceinline__ void load(C2<T> (&a)[N]) const {
#pragma unroll
    for (int r = 0; r < N; r++) {
      const V v = gload<NT>(reinterpret_cast<const V*>(tp + cur->go[r]));
      a[r].x = v.x;
      a[r].y = v.y;
    }
  }
  __device__ __forceinline__ void store(const C2<T> (&a)[N]) const {
#pragma unroll
    for (int r = 0; r < N; r++) {
      V v;
      v.x = a[r].x;
      v.y = a[r].y;
      gstore<NT>(reinterpret_cast<V*>(tp + cur->go[r]), v);
    }
  }
  __device__ __forceinline__ void put(const C2<T> (&a)[N]) const {
#pragma unroll
    for (int r = 0; r < N; r++) lds[ts ^ cur->ls[r]] = a[r];
  }
  __device__ __forceinline__ void get(C2<T> (&a)[N]) const {
#pragma unroll
    for (int r = 0; r < N; r++) a[r] = lds[ts ^ cur->ls[r]];
  }
  // Stage change through a buffer of 2^m scalars (half the LDS of put/get):
  // real parts first, then imaginary parts, each written in the current
  // stage's layout and read in stage `next`'s.  Ends in stage `next`.
  // The caller orders earlier reads of `buf` before this call.
  // Addresses are formed in bytes ((t ^ l) * 8 == t * 8 ^ l * 8) with the
  // register part's byte offsets precomputed (StageDev::lb): one XOR per
  // access.
  __device__ __forceinline__ void exchange_split(C2<T> (&a)[N], uint32_t next, T* __restrict__ buf) {
    static_assert(sizeof(T) == 8, "StageDev::lb holds offsets of 8-byte elements");
    char* const b = reinterpret_cast<char*>(buf);
    const uint32_t ob = ts * (uint32_t)sizeof(T);
    const StageDev* ocur = cur;
#pragma unroll
    for (int r = 0; r < N; r++) *reinterpret_cast<T*>(b + (ob ^ ocur->lb[r])) = a[r].x;
    __syncthreads();
    enter(next);
    const uint32_t nb = ts * (uint32_t)sizeof(T);
#pragma unroll
    for (int r = 0; r < N; r++) a[r].x = *reinterpret_cast<const T*>(b + (nb ^ cur->lb[r]));
    __syncthreads();
#pragma unroll
    for (int r = 0; r < N; r++) *reinterpret_cast<T*>(b + (ob ^ ocur->lb[r])) = a[r].y;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < N; r++) a[r].y = *reinterpret_cast<const T*>(b + (nb ^ cur->lb[r]));
  }
};

}  // namespace fused
}  // namespace aleph
