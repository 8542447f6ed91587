// Host-side gate algebra: matrices, validation and structure tests.
//
// Matrices follow crates/aleph-core/src/gate/kinds.rs:213-399 (row-major,
// qubits[0] the MSB of the index); checks follow the order of
// crates/aleph-sv/src/backend.rs:74-151 (arity, range/duplicates, then the
// UnitaryKq validator or matrix + unitarity ≤ 1e-10).
#pragma once

#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <vector>

#include "common.hpp"

namespace aleph {

using cplx = std::complex<double>;

constexpr double kAmplitudeTol = 1e-10;    // crates/aleph-core/src/lib.rs:53
// Norm² drift budget of FP32 states: a fixed 1e-3, as the reference's GPU
// FP32 readout (crates/aleph-cuda/src/sv/readout_f32.rs:34-39); the FP64
// budget sqrt(2^n) * kAmplitudeTol would reject every legitimate FP32 state.
constexpr double kFp32NormDrift = 1e-3;
constexpr double kDiagonalEpsSq = 1e-30;   // crates/aleph-sv/src/kernels/mod.rs:295
constexpr uint32_t kMaxFusedK = 6;         // widest dense block this backend applies

inline const char* gate_name(uint32_t kind) {
  static const char* names[ALEPH_HIP_GATE_COUNT] = {
      "H", "X", "Y", "Z", "S", "Sdg", "T", "Tdg", "Rx", "Ry", "Rz", "Phase", "U3",
      "Cnot", "Cz", "Swap", "Iswap", "IswapDg", "CRx", "CRy", "CRz", "Toffoli", "Ccz",
      "Unitary1q", "Unitary2q", "UnitaryKq", "Unitary1qDiag", "Unitary2qDiag"};
  return kind < ALEPH_HIP_GATE_COUNT ? names[kind] : "unknown";
}

inline int gate_arity(uint32_t kind, uint32_t k) {
  switch (kind) {
    case ALEPH_HIP_GATE_CNOT: case ALEPH_HIP_GATE_CZ: case ALEPH_HIP_GATE_SWAP:
    case ALEPH_HIP_GATE_ISWAP: case ALEPH_HIP_GATE_ISWAPDG: case ALEPH_HIP_GATE_CRX:
    case ALEPH_HIP_GATE_CRY: case ALEPH_HIP_GATE_CRZ: case ALEPH_HIP_GATE_UNITARY2Q:
    case ALEPH_HIP_GATE_UNITARY2QDIAG:
      return 2;
    case ALEPH_HIP_GATE_TOFFOLI: case ALEPH_HIP_GATE_CCZ:
      return 3;
    case ALEPH_HIP_GATE_UNITARYKQ:
      return (int)k;
    default:
      return kind < ALEPH_HIP_GATE_COUNT ? 1 : -1;
  }
}

inline int gate_param_count(uint32_t kind) {
  switch (kind) {
    case ALEPH_HIP_GATE_RX: case ALEPH_HIP_GATE_RY: case ALEPH_HIP_GATE_RZ:
    case ALEPH_HIP_GATE_PHASE: case ALEPH_HIP_GATE_CRX: case ALEPH_HIP_GATE_CRY:
    case ALEPH_HIP_GATE_CRZ:
      return 1;
    case ALEPH_HIP_GATE_U3:
      return 3;
    default:
      return 0;
  }
}

// Number of reals in the matrix payload of a program record.
inline size_t gate_payload_reals(uint32_t kind, uint32_t k) {
  switch (kind) {
    case ALEPH_HIP_GATE_UNITARY1Q: return 8;
    case ALEPH_HIP_GATE_UNITARY2Q: return 32;
    case ALEPH_HIP_GATE_UNITARYKQ: return 2ull << (2 * k);
    case ALEPH_HIP_GATE_UNITARY1QDIAG: return 4;
    case ALEPH_HIP_GATE_UNITARY2QDIAG: return 8;
    default: return 0;
  }
}

// A validated gate, lowered to its dense matrix (dim x dim, row-major).
struct DenseGate {
  uint32_t kind = 0;
  std::vector<uint32_t> qubits;   // qubits[0] = MSB of the matrix index
  std::vector<uint32_t> controls;
  uint32_t dim = 0;
  std::vector<cplx> m;
  bool diagonal = false;          // every off-diagonal entry has |z|^2 < 1e-30
};

inline std::vector<cplx> gate_matrix(uint32_t kind, const double* p, const double* payload, uint32_t k,
                                     uint32_t* dim_out) {
  const cplx z(0, 0), o(1, 0), i1(0, 1);
  std::vector<cplx> m;
  auto dense = [&](uint32_t d) { m.assign((size_t)d * d, z); *dim_out = d; };
  auto set2 = [&](cplx a, cplx b, cplx c, cplx d) { dense(2); m[0] = a; m[1] = b; m[2] = c; m[3] = d; };
  switch (kind) {
    case ALEPH_HIP_GATE_H: { const double r = 0.70710678118654752440; set2(r, r, r, -r); break; }
    case ALEPH_HIP_GATE_X: set2(z, o, o, z); break;
    case ALEPH_HIP_GATE_Y: set2(z, -i1, i1, z); break;
    case ALEPH_HIP_GATE_Z: set2(o, z, z, -o); break;
    case ALEPH_HIP_GATE_S: set2(o, z, z, i1); break;
    case ALEPH_HIP_GATE_SDG: set2(o, z, z, -i1); break;
    case ALEPH_HIP_GATE_T: set2(o, z, z, cplx(std::cos(M_PI / 4), std::sin(M_PI / 4))); break;
    case ALEPH_HIP_GATE_TDG: set2(o, z, z, cplx(std::cos(M_PI / 4), -std::sin(M_PI / 4))); break;
    case ALEPH_HIP_GATE_RX: {
      const double h = p[0] / 2;
      set2(cplx(std::cos(h), 0), cplx(0, -std::sin(h)), cplx(0, -std::sin(h)), cplx(std::cos(h), 0));
      break;
    }
    case ALEPH_HIP_GATE_RY: {
      const double h = p[0] / 2;
      set2(cplx(std::cos(h), 0), cplx(-std::sin(h), 0), cplx(std::sin(h), 0), cplx(std::cos(h), 0));
      break;
    }
    case ALEPH_HIP_GATE_RZ: {
      const double h = p[0] / 2;
      set2(cplx(std::cos(h), -std::sin(h)), z, z, cplx(std::cos(h), std::sin(h)));
      break;
    }
    case ALEPH_HIP_GATE_PHASE: set2(o, z, z, cplx(std::cos(p[0]), std::sin(p[0]))); break;
    case ALEPH_HIP_GATE_U3: {
      const double c = std::cos(p[0] / 2), s = std::sin(p[0] / 2), phi = p[1], lam = p[2];
      const cplx el(std::cos(lam), std::sin(lam)), ef(std::cos(phi), std::sin(phi));
      const cplx efl(std::cos(phi + lam), std::sin(phi + lam));
      set2(cplx(c, 0), -el * s, ef * s, efl * c);
      break;
    }
    case ALEPH_HIP_GATE_CNOT: dense(4); m[0] = o; m[5] = o; m[2 * 4 + 3] = o; m[3 * 4 + 2] = o; break;
    case ALEPH_HIP_GATE_CZ: dense(4); m[0] = o; m[5] = o; m[10] = o; m[15] = -o; break;
    case ALEPH_HIP_GATE_SWAP: dense(4); m[0] = o; m[1 * 4 + 2] = o; m[2 * 4 + 1] = o; m[15] = o; break;
    case ALEPH_HIP_GATE_ISWAP: dense(4); m[0] = o; m[1 * 4 + 2] = i1; m[2 * 4 + 1] = i1; m[15] = o; break;
    case ALEPH_HIP_GATE_ISWAPDG: dense(4); m[0] = o; m[1 * 4 + 2] = -i1; m[2 * 4 + 1] = -i1; m[15] = o; break;
    case ALEPH_HIP_GATE_CRX: case ALEPH_HIP_GATE_CRY: case ALEPH_HIP_GATE_CRZ: {
      uint32_t d1 = 0;
      const uint32_t base = kind == ALEPH_HIP_GATE_CRX ? ALEPH_HIP_GATE_RX
                            : kind == ALEPH_HIP_GATE_CRY ? ALEPH_HIP_GATE_RY : ALEPH_HIP_GATE_RZ;
      std::vector<cplx> r = gate_matrix(base, p, nullptr, 0, &d1);
      dense(4);
      m[0] = o; m[5] = o;
      m[2 * 4 + 2] = r[0]; m[2 * 4 + 3] = r[1]; m[3 * 4 + 2] = r[2]; m[3 * 4 + 3] = r[3];
      break;
    }
    case ALEPH_HIP_GATE_TOFFOLI:
      dense(8);
      for (int d = 0; d < 6; d++) m[d * 8 + d] = o;
      m[6 * 8 + 7] = o; m[7 * 8 + 6] = o;
      break;
    case ALEPH_HIP_GATE_CCZ:
      dense(8);
      for (int d = 0; d < 7; d++) m[d * 8 + d] = o;
      m[63] = -o;
      break;
    case ALEPH_HIP_GATE_UNITARY1Q: case ALEPH_HIP_GATE_UNITARY2Q: case ALEPH_HIP_GATE_UNITARYKQ: {
      const uint32_t d = kind == ALEPH_HIP_GATE_UNITARY1Q ? 2u : kind == ALEPH_HIP_GATE_UNITARY2Q ? 4u : (1u << k);
      dense(d);
      for (size_t e = 0; e < (size_t)d * d; e++) m[e] = cplx(payload[2 * e], payload[2 * e + 1]);
      break;
    }
    case ALEPH_HIP_GATE_UNITARY1QDIAG: case ALEPH_HIP_GATE_UNITARY2QDIAG: {
      const uint32_t d = kind == ALEPH_HIP_GATE_UNITARY1QDIAG ? 2u : 4u;
      dense(d);
      for (uint32_t e = 0; e < d; e++) m[e * d + e] = cplx(payload[2 * e], payload[2 * e + 1]);
      break;
    }
    default:
      *dim_out = 0;
  }
  return m;
}

inline bool matrix_is_diagonal(const std::vector<cplx>& m, uint32_t dim) {
  for (uint32_t r = 0; r < dim; r++)
    for (uint32_t c = 0; c < dim; c++) {
      if (r == c) continue;
      const cplx e = m[(size_t)r * dim + c];
      if (!std::isfinite(e.real()) || !std::isfinite(e.imag())) return false;
      if (std::norm(e) >= kDiagonalEpsSq) return false;
    }
  return true;
}

// max_{i,j} |(U U†)_{ij} - δ_{ij}|, NaN-propagating (validation.rs:25-55).
inline double unitarity_deviation(const std::vector<cplx>& m, uint32_t dim) {
  double worst = 0.0;
  for (uint32_t i = 0; i < dim; i++)
    for (uint32_t j = 0; j < dim; j++) {
      cplx acc(0, 0);
      for (uint32_t c = 0; c < dim; c++) acc += m[(size_t)i * dim + c] * std::conj(m[(size_t)j * dim + c]);
      const double dev = std::abs(acc - cplx(i == j ? 1.0 : 0.0, 0.0));
      if (std::isnan(dev)) return NAN;
      worst = dev > worst ? dev : worst;
    }
  return worst;
}

inline Failure make_failure(int32_t code, const char* kind, const std::string& msg) {
  Failure f{code, msg};
  f.kind = kind;
  return f;
}

// Validate a GateInstance against an n-qubit state and lower it.
inline DenseGate validate_gate(const aleph_hip_gate& g, uint32_t n) {
  const char* name = gate_name(g.kind);
  const int ar = gate_arity(g.kind, g.k);
  if (ar < 0) throw make_failure(ALEPH_HIP_ERR_UNSUPPORTED_GATE, name, "unknown gate kind");
  if ((uint32_t)ar != g.n_qubits) {
    Failure f = make_failure(ALEPH_HIP_ERR_ARITY_MISMATCH, name, "arity mismatch");
    f.expected = (uint32_t)ar;
    f.got = g.n_qubits;
    throw f;
  }
  uint64_t seen = 0;
  for (uint32_t i = 0; i < g.n_qubits + g.n_controls; i++) {
    const uint32_t q = i < g.n_qubits ? g.qubits[i] : g.controls[i - g.n_qubits];
    if (q >= n) {
      Failure f = make_failure(ALEPH_HIP_ERR_QUBIT_OUT_OF_RANGE, name, "qubit out of range");
      f.qubit = q;
      f.num_qubits = n;
      throw f;
    }
    if (seen & (1ull << q)) {
      Failure f = make_failure(ALEPH_HIP_ERR_DUPLICATE_QUBIT, name, "duplicate qubit");
      f.qubit = q;
      throw f;
    }
    seen |= 1ull << q;
  }
  DenseGate d;
  d.kind = g.kind;
  d.qubits.assign(g.qubits, g.qubits + g.n_qubits);
  d.controls.assign(g.controls, g.controls + g.n_controls);
  if (g.kind == ALEPH_HIP_GATE_UNITARYKQ) {
    // validate_unitary_kq (validation.rs:117-160), width range [2, 6].
    if (g.k < 2 || g.k > kMaxFusedK)
      throw make_failure(ALEPH_HIP_ERR_INVALID_STATE, "UnitaryKq width must be between 2 and 6", "bad width");
    if (g.n_controls != 0) throw make_failure(ALEPH_HIP_ERR_UNSUPPORTED_GATE, "controlled UnitaryKq", "controlled");
    if (g.matrix == nullptr) throw make_failure(ALEPH_HIP_ERR_INVALID_ARGUMENT, name, "missing matrix");
    const size_t w = 1ull << g.k;
    for (size_t e = 0; e < 2 * w * w; e++)
      if (!std::isfinite(g.matrix[e])) throw make_failure(ALEPH_HIP_ERR_NON_FINITE_PARAM, name, "non-finite");
  } else {
    for (int i = 0; i < gate_param_count(g.kind); i++)
      if (!std::isfinite(g.params[i])) throw make_failure(ALEPH_HIP_ERR_NON_FINITE_PARAM, name, "non-finite");
    if (gate_payload_reals(g.kind, g.k) && g.matrix == nullptr)
      throw make_failure(ALEPH_HIP_ERR_INVALID_ARGUMENT, name, "missing matrix");
  }
  d.m = gate_matrix(g.kind, g.params, g.matrix, g.k, &d.dim);
  if (g.kind == ALEPH_HIP_GATE_UNITARYKQ) {
    double worst2 = 0.0;
    for (uint32_t i = 0; i < d.dim; i++)
      for (uint32_t j = 0; j <= i; j++) {
        cplx acc(0, 0);
        for (uint32_t c = 0; c < d.dim; c++) acc += d.m[(size_t)i * d.dim + c] * std::conj(d.m[(size_t)j * d.dim + c]);
        const double e = std::norm(acc - cplx(i == j ? 1.0 : 0.0, 0.0));
        if (!std::isfinite(e)) {
          Failure f = make_failure(ALEPH_HIP_ERR_NON_UNITARY_MATRIX, name, "non-unitary");
          f.value = e;
          throw f;
        }
        worst2 = e > worst2 ? e : worst2;
      }
    if (worst2 > kAmplitudeTol * kAmplitudeTol) {
      Failure f = make_failure(ALEPH_HIP_ERR_NON_UNITARY_MATRIX, name, "non-unitary");
      f.value = std::sqrt(worst2);
      throw f;
    }
  } else {
    const double dev = unitarity_deviation(d.m, d.dim);
    if (!std::isfinite(dev) || dev > kAmplitudeTol) {
      Failure f = make_failure(ALEPH_HIP_ERR_NON_UNITARY_MATRIX, name, "non-unitary");
      f.value = dev;
      throw f;
    }
  }
  d.diagonal = matrix_is_diagonal(d.m, d.dim);
  return d;
}

}  // namespace aleph
