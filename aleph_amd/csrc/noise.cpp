// Host side of the noise trajectories: validation of channels and readout
// errors (crates/aleph-sv/src/noise/execution.rs:207-300) and lowering of a
// decoded program plus its channels into trajectory records.
#include "noise_host.hpp"

#include <cmath>
#include <complex>

#include "common.hpp"
#include "gates.hpp"

namespace aleph {
namespace {

Failure invalid_channel(const std::string& why) {
  Failure f{ALEPH_HIP_ERR_INVALID_ARGUMENT, why};
  f.kind = "invalid-channel";
  return f;
}

void push_complex(std::vector<double>& v, const cplx& z) {
  v.push_back(z.real());
  v.push_back(z.imag());
}

}  // namespace

TrajHost build_trajectories(const std::vector<ProgInstr>& prog, uint32_t n, const aleph_hip_channel* channels,
                            uint32_t n_channels, const double* readout, uint64_t readout_mask) {
  TrajHost h;
  if (n_channels && !channels) throw invalid_channel("null channel array");
  uint32_t next = 0;
  for (uint32_t c = 0; c < n_channels; c++) {
    const aleph_hip_channel& ch = channels[c];
    if (ch.after >= prog.size()) throw invalid_channel("channel attached past the end of the program");
    if (c && ch.after < channels[c - 1].after) throw invalid_channel("channels must be ordered by instruction");
    if (prog[ch.after].type != ProgInstr::GATE) throw invalid_channel("channels attach to gates only");
    if (ch.n_branches == 0) throw invalid_channel("channel without branches");
    for (uint32_t j = 0; j < ch.n_qubits && j < 2; j++)
      if (ch.qubits[j] >= n) throw invalid_channel("channel qubit out of range");
    if (ch.kind == ALEPH_HIP_CHANNEL_PAULI) {
      if (ch.n_qubits < 1 || ch.n_qubits > 2) throw invalid_channel("Pauli channels act on one or two qubits");
      if (ch.n_qubits == 2 && ch.qubits[0] == ch.qubits[1]) throw invalid_channel("duplicate channel qubit");
      if (!ch.weights || !ch.paulis) throw invalid_channel("null Pauli channel data");
      double total = 0.0;
      for (uint32_t b = 0; b < ch.n_branches; b++) {
        const double p = ch.weights[b];
        if (!std::isfinite(p) || p < 0.0) throw invalid_channel("invalid Pauli probability or term arity");
        for (uint32_t j = 0; j < ch.n_qubits; j++)
          if (ch.paulis[b * ch.n_qubits + j] > 3) throw invalid_channel("invalid Pauli code");
        total += p;
      }
      if (!std::isfinite(total) || std::fabs(total - 1.0) > 1e-9)
        throw invalid_channel("Pauli probabilities must sum to one");
    } else if (ch.kind == ALEPH_HIP_CHANNEL_KRAUS) {
      if (ch.n_qubits != 1) throw invalid_channel("Kraus channels act on one qubit");
      if (ch.n_branches > kNoiseMaxKraus) throw invalid_channel("too many Kraus operators");
      if (!ch.kraus) throw invalid_channel("null Kraus channel data");
      cplx comp[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
      for (uint32_t b = 0; b < ch.n_branches; b++) {
        const double* k = ch.kraus + 8 * b;
        for (int e = 0; e < 8; e++)
          if (!std::isfinite(k[e])) throw invalid_channel("non-finite Kraus matrix");
        const cplx K[2][2] = {{cplx(k[0], k[1]), cplx(k[2], k[3])}, {cplx(k[4], k[5]), cplx(k[6], k[7])}};
        for (int i = 0; i < 2; i++)
          for (int j = 0; j < 2; j++)
            for (int r = 0; r < 2; r++) comp[i][j] += std::conj(K[r][i]) * K[r][j];
      }
      for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) {
          const double dev = std::abs(comp[i][j] - cplx(i == j ? 1.0 : 0.0, 0.0));
          if (!std::isfinite(dev) || dev > 1e-9) throw invalid_channel("Kraus operators must be trace preserving");
        }
    } else {
      throw invalid_channel("unknown channel kind");
    }
  }
  // Readout errors: rows of m[t][o] non-negative and summing to one.
  if (n < 64 && (readout_mask >> n) != 0) throw invalid_channel("invalid readout probability or qubit");
  h.readout.assign(2 * (size_t)n, 0.0);
  if (readout_mask) {
    if (!readout) throw invalid_channel("null readout matrices");
    for (uint32_t q = 0; q < n; q++) {
      if (!((readout_mask >> q) & 1ull)) continue;
      const double* m = readout + 4 * q;
      for (int e = 0; e < 4; e++)
        if (!std::isfinite(m[e]) || m[e] < 0.0) throw invalid_channel("invalid readout probability or qubit");
      if (std::fabs(m[0] + m[1] - 1.0) > 1e-9 || std::fabs(m[2] + m[3] - 1.0) > 1e-9)
        throw invalid_channel("invalid readout probability or qubit");
      h.readout[2 * q] = m[1];       // P(read 1 | true 0)
      h.readout[2 * q + 1] = m[2];   // P(read 0 | true 1)
    }
  }
  h.readout_mask = readout_mask;

  for (uint32_t i = 0; i < prog.size(); i++) {
    const ProgInstr& pi = prog[i];
    if (pi.type == ProgInstr::GATE) {
      const DenseGate& d = pi.gate;
      TrajOp op;
      op.k = (uint32_t)d.qubits.size();
      for (uint32_t t = 0; t < op.k; t++) op.q[t] = d.qubits[t];
      for (uint32_t cq : d.controls) op.cmask |= 1ull << cq;
      op.off = (uint32_t)h.reals.size();
      if (d.diagonal) {
        op.kind = TrajOp::DIAG;
        for (uint32_t j = 0; j < d.dim; j++) push_complex(h.reals, d.m[(size_t)j * d.dim + j]);
      } else {
        if (op.k > 3) {
          Failure f{ALEPH_HIP_ERR_INVALID_ARGUMENT, "dense gates on more than 3 qubits are not supported under noise"};
          f.kind = "unsupported";
          throw f;
        }
        op.kind = TrajOp::DENSE;
        for (const cplx& z : d.m) push_complex(h.reals, z);
      }
      h.ops.push_back(op);
    } else if (pi.type == ProgInstr::DIAG_PHASE) {
      TrajOp op;
      op.kind = TrajOp::PHASE;
      op.count = (uint32_t)pi.angles.size();
      op.off = (uint32_t)h.reals.size();
      op.off2 = (uint32_t)h.masks.size();
      op.off3 = (uint32_t)h.index.size();
      uint32_t at = 0;
      h.index.push_back(0);
      for (size_t t = 0; t < pi.angles.size(); t++) {
        h.reals.push_back(pi.angles[t]);
        for (uint64_t c : pi.conds[t]) h.masks.push_back(c);
        at += (uint32_t)pi.conds[t].size();
        h.index.push_back(at);
      }
      if (op.count) h.ops.push_back(op);
    }
    for (; next < n_channels && channels[next].after == i; next++) {
      const aleph_hip_channel& ch = channels[next];
      TrajOp op;
      op.count = ch.n_branches;
      op.off = (uint32_t)h.reals.size();
      op.k = ch.n_qubits;
      op.q[0] = ch.qubits[0];
      op.q[1] = ch.qubits[1];
      if (ch.kind == ALEPH_HIP_CHANNEL_PAULI) {
        op.kind = TrajOp::PAULI;
        op.off2 = (uint32_t)h.masks.size();
        for (uint32_t b = 0; b < ch.n_branches; b++) {
          h.reals.push_back(ch.weights[b]);
          uint64_t xm = 0, zm = 0;
          for (uint32_t j = 0; j < ch.n_qubits; j++) {
            const uint8_t code = ch.paulis[b * ch.n_qubits + j];
            const uint64_t bit = 1ull << ch.qubits[j];
            if (code == 1 || code == 2) xm |= bit;   // X, Y
            if (code == 2 || code == 3) zm |= bit;   // Y, Z
          }
          h.masks.push_back(xm);
          h.masks.push_back(zm);
        }
      } else {
        op.kind = TrajOp::KRAUS;
        for (uint32_t e = 0; e < 8 * ch.n_branches; e++) h.reals.push_back(ch.kraus[e]);
      }
      h.ops.push_back(op);
    }
  }
  if (h.reals.empty()) h.reals.push_back(0.0);
  if (h.masks.empty()) h.masks.push_back(0);
  if (h.index.empty()) h.index.push_back(0);
  return h;
}

}  // namespace aleph
