// CPU oracle for the dense state-vector path — TEST INFRASTRUCTURE ONLY.
//
// A C++ restatement of the reference's CPU FP64 `sv` backend
// (`NaiveSvBackend`, crates/aleph-sv/src/backend.rs) and of the
// readout it exposes.  Only tests/, the smoke check and bench.py's
// cpu_baseline leg load this library, and only as the checker or as the
// timed CPU baseline; the HIP product never links or calls it.
//
// Followed, function by function (reference paths relative to the
// reference checkout):
//   gate matrices        crates/aleph-core/src/gate/kinds.rs:213-399
//   apply_gate checks    crates/aleph-sv/src/backend.rs:74-151
//   unitarity deviation  crates/aleph-sv/src/validation.rs:25-55
//   UnitaryKq validation crates/aleph-sv/src/validation.rs:117-160
//   1q kernels           crates/aleph-sv/src/kernels/aos.rs:26-256, 413-610
//   2q kernels           crates/aleph-sv/src/kernels/aos.rs:1335-1620, 2193, 2946
//   3q kernels           crates/aleph-sv/src/kernels/aos.rs:2999-3110, 3434, 3753
//   structure classifiers crates/aleph-sv/src/kernels/mod.rs:295-570
//   dense k-qubit        crates/aleph-sv/src/kernels/unitary_kq.rs:27-175
//   diagonal phase       crates/aleph-sv/src/kernels/diagonal_phase.rs:26-45, 106-125
//   tiled block          crates/aleph-sv/src/backend.rs:173-260
//   unpermute            crates/aleph-core/src/perm.rs:25-40
//   validate_state       crates/aleph-sv/src/measure.rs:23-39, sampling.rs:14-59
//   measure              crates/aleph-sv/src/measure.rs:41-103
//   sample (alias)       crates/aleph-sv/src/measure.rs:109-150, sampling.rs:61-199
//   expectation          crates/aleph-sv/src/measure.rs:155-235, 289-305
//   probabilities        crates/aleph-sv/src/measure.rs:239-285
//
// Deliberate difference: the reference draws its random numbers from
// rand 0.8.6's StdRng (ChaCha12).  That crate is not present here, so
// the oracle uses xoshiro256** seeded through SplitMix64 and keeps the
// same consumption contract (one u64 per alias draw, one f64 per
// genuine measurement).  Seeded shot lists are therefore not bit-equal
// to the reference's; sampling is checked statistically.
//
// Parity pinning of every function here is done by tests/test_oracle.py
// (Aer fixtures, Kronecker brute force, closed forms, reference KATs).

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <vector>
#include <omp.h>

namespace {

using cd = std::complex<double>;

// Error numbering shared with the product's C ABI (mirrors BackendError,
// crates/aleph-backend/src/lib.rs:22-81).
enum Err : int32_t {
  OK = 0,
  DEVICE_FAILURE = 1,
  QUBIT_OUT_OF_RANGE = 2,
  DUPLICATE_QUBIT = 3,
  ARITY_MISMATCH = 4,
  UNSUPPORTED_GATE = 5,
  UNSUPPORTED_INSTRUCTION = 6,
  SYMBOLIC_PARAM = 7,
  NON_FINITE_PARAM = 8,
  NON_UNITARY_MATRIX = 9,
  EMPTY_CIRCUIT = 10,
  DEGENERATE_MEASUREMENT = 11,
  TOO_MANY_QUBITS = 12,
  INVALID_PAULI_STRING = 13,
  INVALID_STATE = 14,
  INVALID_ARGUMENT = 17,
};

// Gate kinds, in the order of the reference `Gate` enum (kinds.rs:13-115).
enum Kind : uint32_t {
  H = 0, X, Y, Z, S, SDG, T, TDG, RX, RY, RZ, PHASE, U3,
  CNOT, CZ, SWAP, ISWAP, ISWAPDG, CRX, CRY, CRZ, TOFFOLI, CCZ,
  UNITARY1Q, UNITARY2Q, UNITARYKQ, UNITARY1QDIAG, UNITARY2QDIAG,
  NUM_KINDS
};

const double AMPLITUDE_TOL = 1e-10;          // aleph-core/src/lib.rs:53
const double DIAGONAL_EPS_SQ = 1e-30;        // kernels/mod.rs:295
const double PERM_TOL = 1e-14;               // kernels/mod.rs:347
const double DEGENERATE_BRANCH = 1e-300;     // measure.rs:12

struct State {
  uint32_t n;
  std::vector<cd> a;
};

struct ErrInfo {
  int32_t code = OK;
  uint32_t qubit = 0, num_qubits = 0, expected = 0, got = 0;
  double value = 0.0;
  const char* kind = "";
};
thread_local ErrInfo g_err;

int32_t fail(int32_t code) { g_err.code = code; return code; }

int arity_of(uint32_t k, uint32_t kq) {
  if (k <= U3 || k == UNITARY1Q || k == UNITARY1QDIAG) return 1;
  if ((k >= CNOT && k <= CRZ) || k == UNITARY2Q || k == UNITARY2QDIAG) return 2;
  if (k == TOFFOLI || k == CCZ) return 3;
  if (k == UNITARYKQ) return (int)kq;
  return -1;
}

const char* name_of(uint32_t k) {
  static const char* names[] = {
      "H", "X", "Y", "Z", "S", "Sdg", "T", "Tdg", "Rx", "Ry", "Rz", "Phase", "U3",
      "Cnot", "Cz", "Swap", "Iswap", "IswapDg", "CRx", "CRy", "CRz", "Toffoli", "Ccz",
      "Unitary1q", "Unitary2q", "UnitaryKq", "Unitary1qDiag", "Unitary2qDiag"};
  return k < NUM_KINDS ? names[k] : "?";
}

int n_params(uint32_t k) {
  switch (k) {
    case RX: case RY: case RZ: case PHASE: case CRX: case CRY: case CRZ: return 1;
    case U3: return 3;
    default: return 0;
  }
}

// Gate::matrix (kinds.rs:213-399).  Returns dim; fills row-major `m`.
int gate_matrix(uint32_t k, const double* p, const double* mat, std::vector<cd>& m) {
  const cd z(0, 0), o(1, 0), I(0, 1);
  const double r2 = 0.70710678118654752440;  // FRAC_1_SQRT_2
  auto M2 = [&](cd a, cd b, cd c, cd d) { m = {a, b, c, d}; return 2; };
  auto diag4 = [&](cd a, cd b, cd c, cd d) {
    m.assign(16, z); m[0] = a; m[5] = b; m[10] = c; m[15] = d; return 4;
  };
  switch (k) {
    case H: return M2(cd(r2, 0), cd(r2, 0), cd(r2, 0), cd(-r2, 0));
    case X: return M2(z, o, o, z);
    case Y: return M2(z, cd(0, -1), cd(0, 1), z);
    case Z: return M2(o, z, z, -o);
    case S: return M2(o, z, z, I);
    case SDG: return M2(o, z, z, cd(0, -1));
    case T: return M2(o, z, z, cd(std::cos(M_PI / 4), std::sin(M_PI / 4)));
    case TDG: return M2(o, z, z, cd(std::cos(M_PI / 4), -std::sin(M_PI / 4)));
    case RX: { double t = p[0]; cd c(std::cos(t / 2), 0), s(0, -std::sin(t / 2)); return M2(c, s, s, c); }
    case RY: { double t = p[0]; cd c(std::cos(t / 2), 0), s(std::sin(t / 2), 0); return M2(c, -s, s, c); }
    case RZ: { double t = p[0]; return M2(cd(std::cos(t / 2), -std::sin(t / 2)), z, z, cd(std::cos(t / 2), std::sin(t / 2))); }
    case PHASE: return M2(o, z, z, cd(std::cos(p[0]), std::sin(p[0])));
    case U3: {
      double t = p[0], f = p[1], l = p[2];
      cd c(std::cos(t / 2), 0), s(std::sin(t / 2), 0);
      cd el(std::cos(l), std::sin(l)), ef(std::cos(f), std::sin(f)), efl(std::cos(f + l), std::sin(f + l));
      return M2(c, -el * s, ef * s, efl * c);
    }
    case CNOT: m.assign(16, z); m[0] = o; m[5] = o; m[11] = o; m[14] = o; return 4;
    case CZ: return diag4(o, o, o, -o);
    case SWAP: m.assign(16, z); m[0] = o; m[6] = o; m[9] = o; m[15] = o; return 4;
    case ISWAP: m.assign(16, z); m[0] = o; m[6] = I; m[9] = I; m[15] = o; return 4;
    case ISWAPDG: m.assign(16, z); m[0] = o; m[6] = -I; m[9] = -I; m[15] = o; return 4;
    case CRX: { double t = p[0]; cd c(std::cos(t / 2), 0), s(0, -std::sin(t / 2));
      m.assign(16, z); m[0] = o; m[5] = o; m[10] = c; m[11] = s; m[14] = s; m[15] = c; return 4; }
    case CRY: { double t = p[0]; cd c(std::cos(t / 2), 0), s(std::sin(t / 2), 0);
      m.assign(16, z); m[0] = o; m[5] = o; m[10] = c; m[11] = -s; m[14] = s; m[15] = c; return 4; }
    case CRZ: { double t = p[0];
      return diag4(o, o, cd(std::cos(t / 2), -std::sin(t / 2)), cd(std::cos(t / 2), std::sin(t / 2))); }
    case TOFFOLI: m.assign(64, z); for (int i = 0; i < 6; i++) m[i * 9] = o; m[6 * 8 + 7] = o; m[7 * 8 + 6] = o; return 8;
    case CCZ: m.assign(64, z); for (int i = 0; i < 7; i++) m[i * 9] = o; m[63] = -o; return 8;
    case UNITARY1Q: m.resize(4); for (int i = 0; i < 4; i++) m[i] = cd(mat[2 * i], mat[2 * i + 1]); return 2;
    case UNITARY2Q: m.resize(16); for (int i = 0; i < 16; i++) m[i] = cd(mat[2 * i], mat[2 * i + 1]); return 4;
    case UNITARY1QDIAG: return M2(cd(mat[0], mat[1]), z, z, cd(mat[2], mat[3]));
    case UNITARY2QDIAG:
      return diag4(cd(mat[0], mat[1]), cd(mat[2], mat[3]), cd(mat[4], mat[5]), cd(mat[6], mat[7]));
    default: return -1;
  }
}

// validation.rs:25-55 — max |(U U†)_ij − δ_ij|, NaN-propagating.
double unitarity_deviation(const std::vector<cd>& m, int dim) {
  double worst = 0.0;
  for (int i = 0; i < dim; i++)
    for (int j = 0; j < dim; j++) {
      cd acc(0, 0);
      for (int c = 0; c < dim; c++) acc += m[i * dim + c] * std::conj(m[j * dim + c]);
      double dev = std::abs(acc - cd(i == j ? 1.0 : 0.0, 0.0));
      if (std::isnan(dev)) return NAN;
      if (dev > worst) worst = dev;
    }
  return worst;
}

// validation.rs:117-160, with the width range widened to [2, 6].
int32_t validate_kq(uint32_t k, const double* data, uint32_t nc) {
  if (k < 2 || k > 6) { g_err.kind = "UnitaryKq"; return fail(INVALID_STATE); }
  if (nc != 0) { g_err.kind = "controlled UnitaryKq"; return fail(UNSUPPORTED_GATE); }
  const size_t w = 1u << k;
  for (size_t i = 0; i < 2 * w * w; i++)
    if (!std::isfinite(data[i])) { g_err.kind = "UnitaryKq"; return fail(NON_FINITE_PARAM); }
  double dev2 = 0.0;
  for (size_t i = 0; i < w; i++)
    for (size_t j = 0; j <= i; j++) {
      cd s(0, 0);
      for (size_t c = 0; c < w; c++)
        s += cd(data[2 * (i * w + c)], data[2 * (i * w + c) + 1]) *
             std::conj(cd(data[2 * (j * w + c)], data[2 * (j * w + c) + 1]));
      double e = std::norm(s - cd(i == j ? 1.0 : 0.0, 0.0));
      if (!std::isfinite(e)) { g_err.value = e; return fail(NON_UNITARY_MATRIX); }
      dev2 = std::max(dev2, e);
    }
  if (dev2 > AMPLITUDE_TOL * AMPLITUDE_TOL) { g_err.value = std::sqrt(dev2); return fail(NON_UNITARY_MATRIX); }
  return OK;
}

inline uint64_t cmask_of(const uint32_t* c, uint32_t nc) {
  uint64_t m = 0;
  for (uint32_t i = 0; i < nc; i++) m |= 1ull << c[i];
  return m;
}

// ---- kernels (aos.rs scalar bodies; OpenMP over disjoint index groups) ----

bool is_diag(const std::vector<cd>& m, int dim) {  // mod.rs:312-323, 384-400
  for (int r = 0; r < dim; r++)
    for (int c = 0; c < dim; c++) {
      if (r == c) continue;
      cd e = m[r * dim + c];
      if (!std::isfinite(e.real()) || !std::isfinite(e.imag())) return false;
      if (std::norm(e) >= DIAGONAL_EPS_SQ) return false;
    }
  return true;
}

bool is_antidiag2(const std::vector<cd>& m) {  // mod.rs:488-530
  for (int i : {0, 3}) {
    cd e = m[i];
    if (!std::isfinite(e.real()) || !std::isfinite(e.imag())) return false;
    if (std::norm(e) >= DIAGONAL_EPS_SQ) return false;
  }
  return true;
}

void k_1q(State& s, uint32_t t, uint64_t cm, const std::vector<cd>& m) {
  const int64_t len = (int64_t)s.a.size();
  const int64_t tb = 1ll << t;
  cd* a = s.a.data();
  if (is_diag(m, 2)) {  // aos.rs:413-450
    const cd d0 = m[0], d1 = m[3];
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < len; i++)
      if (((uint64_t)i & cm) == cm) a[i] *= (i & tb) ? d1 : d0;
    return;
  }
  if (is_antidiag2(m)) {  // aos.rs:543-610 (X/Y are special cases of the same map)
    const cd u01 = m[1], u10 = m[2];
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < len; i++)
      if (!(i & tb) && ((uint64_t)i & cm) == cm) {
        cd x = a[i], y = a[i | tb];
        a[i] = u01 * y;
        a[i | tb] = u10 * x;
      }
    return;
  }
  const cd m00 = m[0], m01 = m[1], m10 = m[2], m11 = m[3];
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < len; i++)
    if (!(i & tb) && ((uint64_t)i & cm) == cm) {
      cd x = a[i], y = a[i | tb];
      a[i] = m00 * x + m01 * y;
      a[i | tb] = m10 * x + m11 * y;
    }
}

// mod.rs:409-458: canonical 2q permutations.
int classify_perm2(const std::vector<cd>& m) {
  int perm[4];
  for (int r = 0; r < 4; r++) {
    int hit = -1;
    for (int c = 0; c < 4; c++) {
      cd e = m[r * 4 + c];
      double nsq = std::norm(e);
      if (nsq < DIAGONAL_EPS_SQ) continue;
      if (std::abs(nsq - 1.0) < PERM_TOL && std::abs(e.real() - 1.0) < PERM_TOL && std::abs(e.imag()) < PERM_TOL) {
        if (hit >= 0) return -1;
        hit = c;
      } else {
        return -1;
      }
    }
    if (hit < 0) return -1;
    perm[r] = hit;
  }
  bool seen[4] = {false, false, false, false};
  for (int r = 0; r < 4; r++) { if (seen[perm[r]]) return -1; seen[perm[r]] = true; }
  if (perm[0] == 0 && perm[1] == 1 && perm[2] == 2 && perm[3] == 3) return 0;  // identity
  if (perm[0] == 0 && perm[1] == 1 && perm[2] == 3 && perm[3] == 2) return 1;  // CNOT, control = qubits[0]
  if (perm[0] == 0 && perm[1] == 3 && perm[2] == 2 && perm[3] == 1) return 2;  // CNOT, control = qubits[1]
  if (perm[0] == 0 && perm[1] == 2 && perm[2] == 1 && perm[3] == 3) return 3;  // SWAP
  return -1;
}

void k_2q(State& s, uint32_t q0, uint32_t q1, uint64_t cm, const std::vector<cd>& m) {
  const int64_t len = (int64_t)s.a.size();
  const int64_t b0 = 1ll << q0, b1 = 1ll << q1, tm = b0 | b1;
  cd* a = s.a.data();
  int pk = classify_perm2(m);
  if (pk == 0) return;
  if (pk == 1 || pk == 2) {  // aos.rs:1396-1440 (CNOT as a controlled swap)
    const int64_t cb = pk == 1 ? b0 : b1, tb = pk == 1 ? b1 : b0;
    const uint64_t full = cm | (uint64_t)cb;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < len; i++)
      if (!(i & tb) && ((uint64_t)i & full) == full) std::swap(a[i], a[i | tb]);
    return;
  }
  if (pk == 3) {  // aos.rs:1442-1483
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < len; i++)
      if ((i & tm) == b0 && ((uint64_t)i & cm) == cm) std::swap(a[i], a[i ^ tm]);
    return;
  }
  if (is_diag(m, 4)) {  // aos.rs:1485-1568
    const cd d[4] = {m[0], m[5], m[10], m[15]};
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < len; i++)
      if (((uint64_t)i & cm) == cm) a[i] *= d[((i & b0) ? 2 : 0) | ((i & b1) ? 1 : 0)];
    return;
  }
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < len; i++)  // aos.rs:1335-1390
    if (!(i & tm) && ((uint64_t)i & cm) == cm) {
      const int64_t idx[4] = {i, i | b1, i | b0, i | tm};
      cd v[4] = {a[idx[0]], a[idx[1]], a[idx[2]], a[idx[3]]};
      for (int r = 0; r < 4; r++)
        a[idx[r]] = m[r * 4 + 0] * v[0] + m[r * 4 + 1] * v[1] + m[r * 4 + 2] * v[2] + m[r * 4 + 3] * v[3];
    }
}

// Generic dense k-qubit map (unitary_kq.rs:74-175 and aos.rs:3753-3830),
// MSB-first operands, with an external control mask.
void k_kq(State& s, const uint32_t* q, int k, uint64_t cm, const std::vector<cd>& m) {
  const int dim = 1 << k;
  std::vector<int64_t> off(dim);
  for (int r = 0; r < dim; r++) {
    int64_t o = 0;
    for (int p = 0; p < k; p++)
      if ((r >> p) & 1) o |= 1ll << q[k - 1 - p];
    off[r] = o;
  }
  std::vector<uint32_t> sorted(q, q + k);
  std::sort(sorted.begin(), sorted.end());
  const int64_t outer = (int64_t)s.a.size() >> k;
  cd* a = s.a.data();
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < outer; c++) {
    int64_t base = c;
    for (int j = 0; j < k; j++) {  // expand_target_zeros (unitary_kq.rs:102-110)
      int64_t low = (1ll << sorted[j]) - 1;
      base = (base & low) | ((base & ~low) << 1);
    }
    if (((uint64_t)base & cm) != cm) continue;
    cd in[64];
    for (int r = 0; r < dim; r++) in[r] = a[base | off[r]];
    for (int r = 0; r < dim; r++) {
      cd acc(0, 0);
      for (int cc = 0; cc < dim; cc++) acc += m[r * dim + cc] * in[cc];
      a[base | off[r]] = acc;
    }
  }
}

// aos.rs:3434-3460 apply_3q: Toffoli/CCZ specialisations are exact
// permutations/sign flips; the generic path covers every other 8x8.
void k_3q(State& s, const uint32_t* q, uint64_t cm, const std::vector<cd>& m, uint32_t kind) {
  const int64_t len = (int64_t)s.a.size();
  cd* a = s.a.data();
  if (kind == TOFFOLI) {
    const uint64_t full = cm | (1ull << q[0]) | (1ull << q[1]);
    const int64_t tb = 1ll << q[2];
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < len; i++)
      if (!(i & tb) && ((uint64_t)i & full) == full) std::swap(a[i], a[i | tb]);
    return;
  }
  if (kind == CCZ) {
    const uint64_t full = cm | (1ull << q[0]) | (1ull << q[1]) | (1ull << q[2]);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < len; i++)
      if (((uint64_t)i & full) == full) a[i] = -a[i];
    return;
  }
  k_kq(s, q, 3, cm, m);
}

// Tile-major replay of a run of low-target gates (backend.rs:173-260).
// Per tile, gates apply in order; a control at or above the tile width
// is constant over the tile.
void tile_apply_1q(cd* t, int64_t tl, uint32_t q, uint64_t cm, const std::vector<cd>& m) {
  const int64_t tb = 1ll << q;
  for (int64_t i = 0; i < tl; i++)
    if (!(i & tb) && ((uint64_t)i & cm) == cm) {
      cd x = t[i], y = t[i | tb];
      t[i] = m[0] * x + m[1] * y;
      t[i | tb] = m[2] * x + m[3] * y;
    }
}
void tile_apply_2q(cd* t, int64_t tl, uint32_t q0, uint32_t q1, uint64_t cm, const std::vector<cd>& m) {
  const int64_t b0 = 1ll << q0, b1 = 1ll << q1, tm = b0 | b1;
  for (int64_t i = 0; i < tl; i++)
    if (!(i & tm) && ((uint64_t)i & cm) == cm) {
      const int64_t idx[4] = {i, i | b1, i | b0, i | tm};
      cd v[4] = {t[idx[0]], t[idx[1]], t[idx[2]], t[idx[3]]};
      for (int r = 0; r < 4; r++)
        t[idx[r]] = m[r * 4] * v[0] + m[r * 4 + 1] * v[1] + m[r * 4 + 2] * v[2] + m[r * 4 + 3] * v[3];
    }
}

// ---- xoshiro256** (SplitMix64 seeding) -------------------------------------
struct Rng {
  uint64_t s[4];
  explicit Rng(uint64_t seed) {
    for (int i = 0; i < 4; i++) {
      seed += 0x9E3779B97F4A7C15ull;
      uint64_t z = seed;
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
      s[i] = z ^ (z >> 31);
    }
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3]; s[2] ^= t; s[3] = rotl(s[3], 45);
    return r;
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

// measure.rs:23-39 + sampling.rs:14-59.
int32_t validate_state(const State& s, std::vector<double>* probs) {
  if (s.a.empty()) return fail(INVALID_STATE);
  if (probs) probs->resize(s.a.size());
  double total = 0.0;
  for (size_t i = 0; i < s.a.size(); i++) {
    double p = std::norm(s.a[i]);
    if (!std::isfinite(p)) return fail(INVALID_STATE);
    total += p;
    if (probs) (*probs)[i] = p;
  }
  double tol = std::sqrt((double)s.a.size()) * AMPLITUDE_TOL;
  if (!std::isfinite(total) || std::abs(total - 1.0) > tol) return fail(INVALID_STATE);
  return OK;
}

}  // namespace

// ============================================================================
// C ABI (ctypes-friendly).  Every entry point returns an error code; the
// structured detail of the last error is read with orc_last_error.
// ============================================================================
extern "C" {

void orc_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
int orc_get_threads(void) { return omp_get_max_threads(); }

int32_t orc_last_error(uint32_t* qubit, uint32_t* num_qubits, uint32_t* expected, uint32_t* got,
                       double* value, char* kind, int kind_len) {
  if (qubit) *qubit = g_err.qubit;
  if (num_qubits) *num_qubits = g_err.num_qubits;
  if (expected) *expected = g_err.expected;
  if (got) *got = g_err.got;
  if (value) *value = g_err.value;
  if (kind && kind_len > 0) { std::strncpy(kind, g_err.kind, kind_len - 1); kind[kind_len - 1] = 0; }
  return g_err.code;
}

// backend.rs:61-72 (allocate): zeroed, amps[0] = 1, cap check.
int32_t orc_state_new(uint32_t n, uint32_t cap, void** out) {
  g_err = ErrInfo{};
  if (n > cap) { g_err.expected = cap; g_err.got = n; return fail(TOO_MANY_QUBITS); }
  State* s = new State;
  s->n = n;
  s->a.assign(1ull << n, cd(0, 0));
  s->a[0] = cd(1, 0);
  *out = s;
  return OK;
}

void orc_state_free(void* h) { delete (State*)h; }

void orc_state_load(void* h, const double* amps) {
  State* s = (State*)h;
  std::memcpy((void*)s->a.data(), amps, s->a.size() * sizeof(cd));
}

void orc_state_read(void* h, double* out) {
  State* s = (State*)h;
  std::memcpy(out, s->a.data(), s->a.size() * sizeof(cd));
}

// NaiveSvBackend::apply_gate (backend.rs:74-151).  `mat` carries the
// interleaved matrix of the Unitary* kinds (dim*dim complex, or dim
// complex for the *Diag kinds); `k` is the width of UnitaryKq.
int32_t orc_apply_gate(void* h, uint32_t kind, const double* params, const uint32_t* qubits, uint32_t nq,
                       const uint32_t* controls, uint32_t nc, const double* mat, uint32_t k) {
  g_err = ErrInfo{};
  State& s = *(State*)h;
  g_err.kind = name_of(kind);
  int ar = arity_of(kind, k);
  if (ar < 0) return fail(UNSUPPORTED_GATE);
  if ((uint32_t)ar != nq) { g_err.expected = ar; g_err.got = nq; return fail(ARITY_MISMATCH); }
  std::vector<uint32_t> seen;
  for (uint32_t i = 0; i < nq + nc; i++) {
    uint32_t q = i < nq ? qubits[i] : controls[i - nq];
    if (q >= s.n) { g_err.qubit = q; g_err.num_qubits = s.n; return fail(QUBIT_OUT_OF_RANGE); }
    if (std::find(seen.begin(), seen.end(), q) != seen.end()) { g_err.qubit = q; return fail(DUPLICATE_QUBIT); }
    seen.push_back(q);
  }
  const uint64_t cm = cmask_of(controls, nc);
  if (kind == UNITARYKQ) {
    int32_t e = validate_kq(k, mat, nc);
    if (e != OK) return e;
    std::vector<cd> m((size_t)1 << (2 * k));
    for (size_t i = 0; i < m.size(); i++) m[i] = cd(mat[2 * i], mat[2 * i + 1]);
    k_kq(s, qubits, (int)k, 0, m);
    return OK;
  }
  for (int i = 0; i < n_params(kind); i++)
    if (!std::isfinite(params[i])) return fail(NON_FINITE_PARAM);
  std::vector<cd> m;
  int dim = gate_matrix(kind, params, mat, m);
  if (dim < 0) return fail(UNSUPPORTED_GATE);
  double dev = unitarity_deviation(m, dim);
  if (!std::isfinite(dev) || dev > AMPLITUDE_TOL) { g_err.value = dev; return fail(NON_UNITARY_MATRIX); }
  if (dim == 2) k_1q(s, qubits[0], cm, m);
  else if (dim == 4) k_2q(s, qubits[0], qubits[1], cm, m);
  else k_3q(s, qubits, cm, m, kind);
  return OK;
}

// diagonal_phase.rs:26-45, 106-125: ψ[x] *= e^{iφ(x)},
// φ(x) = Σ_t angle_t · [∀ m ∈ conds_t : popcount(m & x) odd].
// CSR layout: term t owns conds[offsets[t] .. offsets[t+1]).
int32_t orc_apply_diagonal_phase(void* h, uint32_t n_terms, const double* angles, const uint64_t* conds,
                                 const uint32_t* offsets) {
  g_err = ErrInfo{};
  State& s = *(State*)h;
  const int64_t len = (int64_t)s.a.size();
  cd* a = s.a.data();
#pragma omp parallel for schedule(static)
  for (int64_t x = 0; x < len; x++) {
    double phi = 0.0;
    for (uint32_t t = 0; t < n_terms; t++) {
      bool all = true;
      for (uint32_t c = offsets[t]; c < offsets[t + 1]; c++)
        if ((__builtin_popcountll(conds[c] & (uint64_t)x) & 1) == 0) { all = false; break; }
      if (all) phi += angles[t];
    }
    const double co = std::cos(phi), si = std::sin(phi);
    const double re = a[x].real() * co - a[x].imag() * si;
    const double im = a[x].real() * si + a[x].imag() * co;
    a[x] = cd(re, im);
  }
  return OK;
}

// backend.rs:173-260: gates (arity <= 2, targets < tile_bits) applied
// tile-major.  Gate records: kinds[g], params[3g..], qubits[2g..],
// controls CSR (ctrl_off[g]..ctrl_off[g+1]), mats[8*4*g..] (4x4 slot).
int32_t orc_apply_tiled_block(void* h, uint32_t tile_bits, uint32_t n_gates, const uint32_t* kinds,
                              const double* params, const uint32_t* qubits, const uint32_t* ctrl_off,
                              const uint32_t* ctrls, const double* mats) {
  g_err = ErrInfo{};
  State& s = *(State*)h;
  std::vector<std::vector<cd>> ms(n_gates);
  std::vector<int> dims(n_gates);
  for (uint32_t g = 0; g < n_gates; g++) {
    dims[g] = gate_matrix(kinds[g], params + 3 * g, mats + 32 * g, ms[g]);
    if (dims[g] != 2 && dims[g] != 4) { g_err.kind = name_of(kinds[g]); return fail(UNSUPPORTED_GATE); }
  }
  const uint32_t n = s.n;
  const uint32_t t = std::min(tile_bits, n);
  const int64_t tl = 1ll << t, n_tiles = (int64_t)s.a.size() >> t;
  cd* a = s.a.data();
#pragma omp parallel for schedule(static)
  for (int64_t ti = 0; ti < n_tiles; ti++) {
    const uint64_t base = (uint64_t)ti << t;
    cd* tile = a + base;
    for (uint32_t g = 0; g < n_gates; g++) {
      bool skip = false;
      uint64_t low = 0;
      for (uint32_t c = ctrl_off[g]; c < ctrl_off[g + 1]; c++) {
        if (ctrls[c] >= t) { if (!((base >> ctrls[c]) & 1)) { skip = true; break; } }
        else low |= 1ull << ctrls[c];
      }
      if (skip) continue;
      if (dims[g] == 2) tile_apply_1q(tile, tl, qubits[2 * g], low, ms[g]);
      else tile_apply_2q(tile, tl, qubits[2 * g], qubits[2 * g + 1], low, ms[g]);
    }
  }
  return OK;
}

// perm.rs:25-40: out[i] = phys[j], j sets bit perm[l] for every set bit l of i.
void orc_unpermute(void* h, const uint32_t* perm) {
  State& s = *(State*)h;
  std::vector<cd> out(s.a.size());
  const int64_t len = (int64_t)s.a.size();
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < len; i++) {
    uint64_t j = 0;
    for (uint32_t l = 0; l < s.n; l++)
      if ((i >> l) & 1) j |= 1ull << perm[l];
    out[i] = s.a[j];
  }
  s.a.swap(out);
}

int32_t orc_validate_state(void* h) { g_err = ErrInfo{}; return validate_state(*(State*)h, nullptr); }

double orc_norm_sqr(void* h) {
  State& s = *(State*)h;
  double t = 0;
  for (auto& z : s.a) t += std::norm(z);
  return t;
}

// measure.rs:41-103.  `u` is the uniform the caller drew; it is used
// only when both branches are non-degenerate (*consumed set to 1).
int32_t orc_measure(void* h, uint32_t qubit, double u, int32_t* outcome, int32_t* consumed) {
  g_err = ErrInfo{};
  State& s = *(State*)h;
  *consumed = 0;
  if (qubit >= s.n) { g_err.qubit = qubit; g_err.num_qubits = s.n; return fail(QUBIT_OUT_OF_RANGE); }
  std::vector<double> probs;
  int32_t e = validate_state(s, &probs);
  if (e != OK) return e;
  const uint64_t qb = 1ull << qubit;
  double p0 = 0, p1 = 0;
  for (size_t i = 0; i < probs.size(); i++) (i & qb) ? p1 += probs[i] : p0 += probs[i];
  p0 = std::min(std::max(p0, 0.0), 1.0);
  p1 = std::min(std::max(p1, 0.0), 1.0);
  bool one_deg = p1 < DEGENERATE_BRANCH, zero_deg = p0 < DEGENERATE_BRANCH;
  bool out;
  double p;
  if (zero_deg && one_deg) { g_err.qubit = qubit; g_err.value = std::max(p0, p1); return fail(DEGENERATE_MEASUREMENT); }
  if (zero_deg) { out = true; p = p1; }
  else if (one_deg) { out = false; p = p0; }
  else { *consumed = 1; out = u < p1; p = out ? p1 : p0; }
  const double nrm = std::sqrt(p);
  for (size_t i = 0; i < s.a.size(); i++) {
    if ((bool)(i & qb) == out) s.a[i] /= cd(nrm, 0);
    else s.a[i] = cd(0, 0);
  }
  *outcome = out;
  return OK;
}

// measure.rs:109-150 + sampling.rs:61-199 (Vose alias for shots >= 2,
// validated inverse CDF for one shot).
int32_t orc_sample(void* h, uint32_t shots, uint64_t seed, uint64_t* out) {
  g_err = ErrInfo{};
  State& s = *(State*)h;
  Rng rng(seed);
  std::vector<double> p;
  int32_t e = validate_state(s, &p);
  if (e != OK) return e;
  if (shots == 0) return OK;
  const size_t n = p.size();
  if (shots == 1) {  // sampling.rs:61-91
    double total = 0;
    for (double x : p) total += x;
    double target = rng.uniform() * total, cum = 0;
    int64_t last = -1;
    for (size_t i = 0; i < n; i++)
      if (p[i] > 0) { last = (int64_t)i; cum += p[i]; if (target < cum) { out[0] = i; return OK; } }
    if (last < 0) return fail(INVALID_STATE);
    out[0] = (uint64_t)last;
    return OK;
  }
  std::vector<double> prob(n, 0.0);
  std::vector<uint32_t> alias(n, 0);
  if (n == 1) { prob[0] = 1.0; }
  else {
    std::vector<double> sc(n);
    std::vector<uint32_t> small, large;
    small.reserve(n); large.reserve(n);
    for (size_t i = 0; i < n; i++) {
      sc[i] = (double)n * p[i];
      (sc[i] < 1.0 ? small : large).push_back((uint32_t)i);
    }
    while (!small.empty() && !large.empty()) {
      uint32_t sm = small.back(); small.pop_back();
      uint32_t lg = large.back(); large.pop_back();
      prob[sm] = sc[sm];
      alias[sm] = lg;
      double nl = (sc[lg] + sc[sm]) - 1.0;
      sc[lg] = nl;
      (nl < 1.0 ? small : large).push_back(lg);
    }
    for (uint32_t i : large) { prob[i] = 1.0; alias[i] = i; }
    for (uint32_t i : small) { prob[i] = 1.0; alias[i] = i; }
  }
  for (uint32_t k = 0; k < shots; k++) {  // sampling.rs:183-199
    uint64_t bits = rng.next();
    size_t i = (size_t)(bits >> 32) & (n - 1);
    double u = (double)(uint32_t)bits * (1.0 / 4294967296.0);
    out[k] = u < prob[i] ? i : alias[i];
  }
  return OK;
}

// measure.rs:155-235.  paulis[i] in {0:I, 1:X, 2:Y, 3:Z}.
int32_t orc_expectation(void* h, double coeff, const uint32_t* qubits, const uint8_t* paulis, uint32_t nt,
                        double* out) {
  g_err = ErrInfo{};
  State& s = *(State*)h;
  std::vector<double> probs;
  int32_t e = validate_state(s, &probs);
  if (e != OK) return e;
  if (!std::isfinite(coeff)) { g_err.kind = "non-finite coefficient"; return fail(INVALID_PAULI_STRING); }
  std::vector<uint32_t> seen;
  for (uint32_t i = 0; i < nt; i++) {
    uint32_t q = qubits[i];
    if (q >= s.n) { g_err.qubit = q; g_err.num_qubits = s.n; return fail(QUBIT_OUT_OF_RANGE); }
    if (std::find(seen.begin(), seen.end(), q) != seen.end()) { g_err.qubit = q; return fail(DUPLICATE_QUBIT); }
    seen.push_back(q);
  }
  uint64_t zm = 0;
  bool all_z = true;
  for (uint32_t i = 0; i < nt; i++) {
    if (paulis[i] == 3) zm |= 1ull << qubits[i];
    else if (paulis[i] != 0) all_z = false;
  }
  if (all_z) {  // measure.rs:289-305
    double acc = 0;
    for (size_t i = 0; i < probs.size(); i++) acc += (__builtin_popcountll(i & zm) & 1) ? -probs[i] : probs[i];
    *out = coeff * acc;
    return OK;
  }
  State tmp = s;
  for (uint32_t i = 0; i < nt; i++) {
    const cd z(0, 0), o(1, 0);
    std::vector<cd> m;
    switch (paulis[i]) {
      case 1: m = {z, o, o, z}; break;
      case 2: m = {z, cd(0, -1), cd(0, 1), z}; break;
      case 3: m = {o, z, z, -o}; break;
      default: continue;
    }
    k_1q(tmp, qubits[i], 0, m);
  }
  cd acc(0, 0);
  for (size_t i = 0; i < s.a.size(); i++) acc += std::conj(s.a[i]) * tmp.a[i];
  *out = coeff * acc.real();
  return OK;
}

// measure.rs:239-285.  Output indexed with qubits[0] as the LSB.
int32_t orc_probabilities(void* h, const uint32_t* qubits, uint32_t nq, double* out) {
  g_err = ErrInfo{};
  State& s = *(State*)h;
  std::vector<double> probs;
  int32_t e = validate_state(s, &probs);
  if (e != OK) return e;
  std::vector<uint32_t> seen;
  for (uint32_t i = 0; i < nq; i++) {
    uint32_t q = qubits[i];
    if (q >= s.n) { g_err.qubit = q; g_err.num_qubits = s.n; return fail(QUBIT_OUT_OF_RANGE); }
    if (std::find(seen.begin(), seen.end(), q) != seen.end()) { g_err.qubit = q; return fail(DUPLICATE_QUBIT); }
    seen.push_back(q);
  }
  if (nq == 0) { out[0] = 1.0; return OK; }
  const size_t od = 1ull << nq;
  std::fill(out, out + od, 0.0);
  for (size_t i = 0; i < probs.size(); i++) {
    size_t k = 0;
    for (uint32_t p = 0; p < nq; p++)
      if ((i >> qubits[p]) & 1) k |= 1ull << p;
    out[k] += probs[i];
  }
  return OK;
}

// Pure gate matrix query (kinds.rs:213-399) so tests can pin the table.
int32_t orc_gate_matrix(uint32_t kind, const double* params, const double* mat, double* out, int32_t* dim) {
  std::vector<cd> m;
  int d = gate_matrix(kind, params, mat, m);
  *dim = d;
  if (d < 0) return UNSUPPORTED_GATE;
  std::memcpy(out, m.data(), m.size() * sizeof(cd));
  return OK;
}

}  // extern "C"
