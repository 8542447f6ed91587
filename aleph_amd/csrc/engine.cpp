// Program decoding, gate lowering, the rule-based pass planner and plan
// execution for the gfx950 state-vector backend.
//
// The planner replaces the reference's per-gate execution
// (crates/aleph-cuda/src/execution.rs:78-131 + sv/backend.rs:515-581, one
// launch per gate or per <=3 disjoint 1q gates) with full-state passes that
// each apply as many gates as a dependency-respecting greedy schedule fits
// into one chunk of m qubits, under a per-pass arithmetic budget.
#include "engine.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>

#include "jit.hpp"
#include "kernels.hpp"

namespace aleph {

namespace {

inline uint64_t qmask(const std::vector<uint32_t>& qs) {
  uint64_t m = 0;
  for (uint32_t q : qs) m |= 1ull << q;
  return m;
}

inline int popc(uint64_t x) { return __builtin_popcountll(x); }

Failure bad_program(const char* why) {
  Failure f{ALEPH_HIP_ERR_INVALID_ARGUMENT, std::string("malformed program: ") + why};
  f.kind = "program";
  return f;
}

}  // namespace

std::vector<ProgInstr> decode_program(const uint32_t* words, size_t n_words, const double* reals, size_t n_reals,
                                      uint32_t n) {
  std::vector<ProgInstr> out;
  size_t w = 0, r = 0;
  auto need_w = [&](size_t c) { if (w + c > n_words) throw bad_program("word stream truncated"); };
  auto need_r = [&](size_t c) { if (r + c > n_reals) throw bad_program("real stream truncated"); };
  while (w < n_words) {
    const uint32_t tag = words[w];
    if (tag == ALEPH_HIP_INSTR_BARRIER) {
      need_w(2);
      const uint32_t cnt = words[w + 1];
      need_w(2 + (size_t)cnt);
      for (uint32_t i = 0; i < cnt; i++)
        if (words[w + 2 + i] >= n) {
          Failure f = make_failure(ALEPH_HIP_ERR_QUBIT_OUT_OF_RANGE, "barrier", "qubit out of range");
          f.qubit = words[w + 2 + i];
          f.num_qubits = n;
          throw f;
        }
      w += 2 + cnt;
      ProgInstr pi;
      pi.type = ProgInstr::BARRIER;
      out.push_back(std::move(pi));
      continue;
    }
    if (tag == ALEPH_HIP_INSTR_DIAGONAL_PHASE) {
      need_w(2);
      const uint32_t nt = words[w + 1];
      w += 2;
      ProgInstr pi;
      pi.type = ProgInstr::DIAG_PHASE;
      need_r(nt);
      for (uint32_t t = 0; t < nt; t++) {
        need_w(1);
        const uint32_t nc = words[w++];
        need_w(2 * (size_t)nc);
        std::vector<uint64_t> cs;
        for (uint32_t c = 0; c < nc; c++) {
          const uint64_t v = (uint64_t)words[w] | ((uint64_t)words[w + 1] << 32);
          w += 2;
          if (n < 64 && (v >> n) != 0) {
            Failure f = make_failure(ALEPH_HIP_ERR_QUBIT_OUT_OF_RANGE, "DiagonalPhase", "mask beyond state");
            f.qubit = 63 - __builtin_clzll(v);
            f.num_qubits = n;
            throw f;
          }
          cs.push_back(v);
        }
        pi.conds.push_back(std::move(cs));
        const double a = reals[r++];
        if (!std::isfinite(a)) throw make_failure(ALEPH_HIP_ERR_NON_FINITE_PARAM, "DiagonalPhase", "non-finite angle");
        pi.angles.push_back(a);
      }
      out.push_back(std::move(pi));
      continue;
    }
    need_w(4);
    aleph_hip_gate g{};
    g.kind = tag;
    g.n_qubits = words[w + 1];
    g.n_controls = words[w + 2];
    g.k = words[w + 3];
    need_w(4 + (size_t)g.n_qubits + g.n_controls);
    g.qubits = words + w + 4;
    g.controls = words + w + 4 + g.n_qubits;
    w += 4 + (size_t)g.n_qubits + g.n_controls;
    if (g.kind >= ALEPH_HIP_GATE_COUNT) throw make_failure(ALEPH_HIP_ERR_UNSUPPORTED_GATE, "unknown", "unknown kind");
    if (g.kind == ALEPH_HIP_GATE_UNITARYKQ && (g.k < 1 || g.k > 16)) throw bad_program("UnitaryKq width");
    const int np = gate_param_count(g.kind);
    need_r((size_t)np);
    for (int i = 0; i < np; i++) g.params[i] = reals[r + i];
    r += np;
    const size_t pr = gate_payload_reals(g.kind, g.k);
    need_r(pr);
    g.matrix = pr ? reals + r : nullptr;
    r += pr;
    ProgInstr pi;
    pi.type = ProgInstr::GATE;
    pi.gate = validate_gate(g, n);
    out.push_back(std::move(pi));
  }
  if (r != n_reals) throw bad_program("unused reals");
  return out;
}

namespace {

// Decompose a diagonal whose entries are unit-modulus into AND-mask phase
// terms: alpha(x) = sum_S beta_S * prod_{q in S} x_q (Moebius inversion over
// subsets of the target bits), each term also requiring every control.
void diag_to_terms(const DenseGate& d, uint64_t cm, std::vector<std::pair<uint64_t, double>>& terms) {
  const uint32_t k = (uint32_t)d.qubits.size();
  const uint32_t dim = 1u << k;
  std::vector<double> alpha(dim);
  for (uint32_t j = 0; j < dim; j++) {
    const cplx e = d.m[(size_t)j * dim + j];
    alpha[j] = std::atan2(e.imag(), e.real());
  }
  for (uint32_t S = 0; S < dim; S++) {
    double beta = 0.0;
    for (uint32_t T = S;; T = (T - 1) & S) {
      const int sign = (popc(S ^ T) & 1) ? -1 : 1;
      beta += sign * alpha[T];
      if (T == 0) break;
    }
    if (beta == 0.0) continue;
    uint64_t mask = cm;
    for (uint32_t p = 0; p < k; p++)
      if ((S >> p) & 1u) mask |= 1ull << d.qubits[k - 1 - p];
    terms.emplace_back(mask, beta);
  }
}

// Expand one parity term (fires when popcount(c & x) is odd for every c in
// conds) whose support has at most kMaxParitySupport bits into AND-mask
// monomials: the indicator is a multilinear polynomial in the support bits,
// and its coefficients are the Moebius transform of its values on subsets.
// An empty condition list fires everywhere (a constant phase, mask 0).
constexpr int kMaxParitySupport = 6;

bool parity_term_to_and(double angle, const std::vector<uint64_t>& conds, std::map<uint64_t, double>& acc) {
  uint64_t support = 0;
  for (uint64_t c : conds) support |= c;
  const int s = popc(support);
  if (s > kMaxParitySupport) return false;
  uint32_t pos[kMaxParitySupport];
  int k = 0;
  for (uint64_t rest = support; rest; rest &= rest - 1) pos[k++] = (uint32_t)__builtin_ctzll(rest);
  const uint32_t dim = 1u << s;
  auto deposit = [&](uint32_t u) {
    uint64_t x = 0;
    for (int i = 0; i < s; i++)
      if ((u >> i) & 1u) x |= 1ull << pos[i];
    return x;
  };
  double a[1 << kMaxParitySupport];
  for (uint32_t u = 0; u < dim; u++) {
    const uint64_t x = deposit(u);
    bool fires = true;
    for (uint64_t c : conds)
      if ((popc(c & x) & 1) == 0) fires = false;
    a[u] = fires ? 1.0 : 0.0;
  }
  for (int i = 0; i < s; i++)
    for (uint32_t u = 0; u < dim; u++)
      if ((u >> i) & 1u) a[u] -= a[u ^ (1u << i)];
  for (uint32_t u = 0; u < dim; u++)
    if (a[u] != 0.0) acc[deposit(u)] += angle * a[u];
  return true;
}

LGate lower_gate(const DenseGate& d) {
  LGate g;
  const uint64_t cm = qmask(d.controls);
  const uint32_t k = (uint32_t)d.qubits.size();
  const uint32_t dim = d.dim;
  g.cmask = cm;
  if (d.diagonal) {
    bool sign = true;
    for (uint32_t j = 0; j < dim; j++) {
      const cplx e = d.m[(size_t)j * dim + j];
      const double want = j + 1 == dim ? -1.0 : 1.0;
      if (e.real() != want || e.imag() != 0.0) sign = false;
    }
    g.dg = d.qubits;
    g.dg.insert(g.dg.end(), d.controls.begin(), d.controls.end());
    if (sign) {
      g.kind = LGate::SIGN;
      g.sign_mask = cm | qmask(d.qubits);
      g.cost = 0.2;
      return g;
    }
    bool unit = true;
    for (uint32_t j = 0; j < dim; j++)
      if (std::fabs(std::abs(d.m[(size_t)j * dim + j]) - 1.0) > 1e-14) unit = false;
    if (unit && k <= 6) {
      g.kind = LGate::TERMS;
      diag_to_terms(d, cm, g.terms);
      g.cost = 0.05 * (double)g.terms.size();
      return g;
    }
    if (k <= 2) {
      g.kind = LGate::DIAGX;
      g.qa = d.qubits[0];
      g.qb = k == 2 ? d.qubits[1] : 0;
      g.n_diag_targets = k;
      for (uint32_t sel = 0; sel < 4; sel++) {
        const uint32_t j = k == 2 ? sel : (sel >> 1);
        const cplx e = d.m[(size_t)j * dim + j];
        g.m[2 * sel] = e.real();
        g.m[2 * sel + 1] = e.imag();
      }
      g.cost = 1.0;
      return g;
    }
    // wide non-unit diagonal: dense standalone sweep
  } else {
    // Controlled structure by kind: the control qubits act diagonally.
    uint32_t tgt = UINT32_MAX;
    uint64_t extra = 0;
    std::vector<cplx> blk;
    if (d.kind == ALEPH_HIP_GATE_CNOT) {
      tgt = d.qubits[1];
      extra = 1ull << d.qubits[0];
      blk = {0, 1, 1, 0};
    } else if (d.kind == ALEPH_HIP_GATE_TOFFOLI) {
      tgt = d.qubits[2];
      extra = (1ull << d.qubits[0]) | (1ull << d.qubits[1]);
      blk = {0, 1, 1, 0};
    } else if (d.kind == ALEPH_HIP_GATE_CRX || d.kind == ALEPH_HIP_GATE_CRY) {
      tgt = d.qubits[1];
      extra = 1ull << d.qubits[0];
      blk = {d.m[10], d.m[11], d.m[14], d.m[15]};
    }
    if (tgt != UINT32_MAX) {
      g.kind = LGate::G1Q;
      g.qa = tgt;
      g.cmask = cm | extra;
      g.nd = {tgt};
      for (uint32_t q = 0; q < 64; q++)
        if ((g.cmask >> q) & 1ull) g.dg.push_back(q);
      for (int e = 0; e < 4; e++) {
        g.m[2 * e] = blk[e].real();
        g.m[2 * e + 1] = blk[e].imag();
      }
      g.cost = 1.0;
      return g;
    }
    if (dim == 2 || dim == 4) {
      g.kind = dim == 2 ? LGate::G1Q : LGate::G2Q;
      g.qa = d.qubits[0];
      g.qb = dim == 4 ? d.qubits[1] : 0;
      g.nd = d.qubits;
      g.dg = d.controls;
      for (uint32_t e = 0; e < dim * dim; e++) {
        g.m[2 * e] = d.m[e].real();
        g.m[2 * e + 1] = d.m[e].imag();
      }
      g.cost = dim == 2 ? 1.0 : 2.0;
      return g;
    }
  }
  g.kind = LGate::KQ;
  g.k = k;
  g.kq_qubits = d.qubits;
  g.nd = d.qubits;
  g.dg = d.controls;
  g.kq_mat.resize(2 * (size_t)dim * dim);
  for (size_t e = 0; e < (size_t)dim * dim; e++) {
    g.kq_mat[2 * e] = d.m[e].real();
    g.kq_mat[2 * e + 1] = d.m[e].imag();
  }
  g.cost = 1e9;
  return g;
}

}  // namespace

std::vector<LGate> lower_program(const std::vector<ProgInstr>& prog, uint32_t n) {
  std::vector<LGate> out;
  for (const ProgInstr& pi : prog) {
    if (pi.type == ProgInstr::BARRIER) continue;
    if (pi.type == ProgInstr::GATE) {
      LGate g = lower_gate(pi.gate);
      const DenseGate& d = pi.gate;
      if (d.controls.empty() && (d.dim == 2 || d.dim == 4)) {
        g.dense_k = d.dim == 2 ? 1 : 2;
        g.dense_q[0] = d.qubits[0];
        g.dense_q[1] = d.dim == 4 ? d.qubits[1] : 0;
        for (uint32_t e = 0; e < d.dim * d.dim; e++) {
          g.dense[2 * e] = d.m[e].real();
          g.dense[2 * e + 1] = d.m[e].imag();
        }
      }
      out.push_back(std::move(g));
      continue;
    }
    // DiagonalPhase: terms on few bits become AND-mask terms for the fused
    // passes; the rest stay a standalone parity-polynomial sweep.  Both
    // parts are diagonal, so their order does not matter.
    std::map<uint64_t, double> acc;
    LGate poly;
    poly.kind = LGate::POLY;
    poly.poly_off.push_back(0);
    uint64_t and_support = 0, poly_support = 0;
    for (size_t t = 0; t < pi.angles.size(); t++) {
      if (parity_term_to_and(pi.angles[t], pi.conds[t], acc)) {
        for (uint64_t c : pi.conds[t]) and_support |= c;
        continue;
      }
      for (uint64_t c : pi.conds[t]) {
        poly.poly_conds.push_back(c);
        poly_support |= c;
      }
      poly.poly_off.push_back((uint32_t)poly.poly_conds.size());
      poly.poly_angles.push_back(pi.angles[t]);
    }
    auto support_list = [&](uint64_t sup) {
      std::vector<uint32_t> v;
      for (uint32_t q = 0; q < 64 && q < n; q++)
        if ((sup >> q) & 1ull) v.push_back(q);
      return v;
    };
    LGate g;
    g.kind = LGate::TERMS;
    for (const auto& kv : acc)
      if (kv.second != 0.0) g.terms.emplace_back(kv.first, kv.second);
    if (!g.terms.empty()) {
      g.dg = support_list(and_support);
      g.cost = 0.05 * (double)g.terms.size();
      out.push_back(std::move(g));
    }
    if (!poly.poly_angles.empty()) {
      poly.dg = support_list(poly_support);
      poly.cost = 1e9;
      out.push_back(std::move(poly));
    }
  }
  return out;
}

namespace {

// Merge runs of uncontrolled 1q gates on one qubit (the effect of the
// reference's Fuse1qRuns, crates/aleph-ir/src/passes/fuse_1q.rs): M = G_k ... G_1.
std::vector<LGate> fuse_1q_runs(const std::vector<LGate>& in) {
  std::vector<LGate> out;
  std::vector<int64_t> open(64, -1);  // index in `out` of the open 1q run per qubit
  for (const LGate& g : in) {
    if (g.kind == LGate::G1Q && g.cmask == 0) {
      const uint32_t q = g.qa;
      if (open[q] >= 0) {
        LGate& acc = out[(size_t)open[q]];
        double r[8];
        const double* a = g.m;     // new gate (left factor)
        const double* b = acc.m;   // accumulated product
        for (int i = 0; i < 2; i++)
          for (int j = 0; j < 2; j++) {
            double re = 0, im = 0;
            for (int c = 0; c < 2; c++) {
              const double ar = a[2 * (i * 2 + c)], ai = a[2 * (i * 2 + c) + 1];
              const double br = b[2 * (c * 2 + j)], bi = b[2 * (c * 2 + j) + 1];
              re += ar * br - ai * bi;
              im += ar * bi + ai * br;
            }
            r[2 * (i * 2 + j)] = re;
            r[2 * (i * 2 + j) + 1] = im;
          }
        std::memcpy(acc.m, r, sizeof(r));
        std::memcpy(acc.dense, r, sizeof(r));
        continue;
      }
      open[q] = (int64_t)out.size();
      out.push_back(g);
      continue;
    }
    for (uint32_t q : g.nd) open[q] = -1;
    for (uint32_t q : g.dg) open[q] = -1;
    out.push_back(g);
  }
  return out;
}

// 4x4 complex helpers (interleaved, row-major, index bit 1 = MSB qubit).
void mul4(const double* a, const double* b, double* out) {
  double t[32];
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) {
      double re = 0, im = 0;
      for (int k = 0; k < 4; k++) {
        const double ar = a[2 * (i * 4 + k)], ai = a[2 * (i * 4 + k) + 1];
        const double br = b[2 * (k * 4 + j)], bi = b[2 * (k * 4 + j) + 1];
        re += ar * br - ai * bi;
        im += ar * bi + ai * br;
      }
      t[2 * (i * 4 + j)] = re;
      t[2 * (i * 4 + j) + 1] = im;
    }
  std::memcpy(out, t, sizeof(t));
}

// Lift a 2x2 on qubit q into the 4x4 of the pair (msb, lsb).
void lift2(const double* m2, bool on_msb, double* out) {
  std::memset(out, 0, 32 * sizeof(double));
  for (int r = 0; r < 4; r++)
    for (int c = 0; c < 4; c++) {
      const int hr = r >> 1, lr = r & 1, hc = c >> 1, lc = c & 1;
      int e = -1;
      if (on_msb && lr == lc) e = hr * 2 + hc;
      if (!on_msb && hr == hc) e = lr * 2 + lc;
      if (e >= 0) {
        out[2 * (r * 4 + c)] = m2[2 * e];
        out[2 * (r * 4 + c) + 1] = m2[2 * e + 1];
      }
    }
}

// Pair fusion: a 2-qubit gate with at least one pending non-diagonal
// 1-qubit gate on its qubits absorbs every pending 1-qubit gate before it
// and every 1-qubit gate after it on either qubit up to that qubit's next
// other gate, into one dense 4x4.  2-qubit gates without pending neighbours
// stay as they are (a bare CZ stays a sign flip).  Moving the absorbed
// gates to the 2-qubit gate's position is valid because no other gate
// touches their qubit in between.
std::vector<LGate> fuse_pairs(const std::vector<LGate>& in) {
  const size_t N = in.size();
  std::vector<int64_t> owner(N, -1);
  std::vector<char> root(N, 0);
  std::vector<std::vector<size_t>> pre(N), post(N);
  std::vector<std::vector<size_t>> pending(64);
  auto touches = [](const LGate& g, uint32_t q) {
    for (uint32_t x : g.nd)
      if (x == q) return true;
    for (uint32_t x : g.dg)
      if (x == q) return true;
    return false;
  };
  for (size_t i = 0; i < N; i++) {
    if (owner[i] >= 0) continue;
    const LGate& g = in[i];
    if (g.dense_k == 1) {
      pending[g.dense_q[0]].push_back(i);
      continue;
    }
    if (g.dense_k == 2 && (g.kind == LGate::G2Q || g.kind == LGate::SIGN || g.kind == LGate::G1Q)) {
      const uint32_t a = g.dense_q[0], b = g.dense_q[1];
      bool any_nd = false;
      for (uint32_t q : {a, b})
        for (size_t j : pending[q]) any_nd |= in[j].kind == LGate::G1Q;
      if (any_nd) {
        root[i] = 1;
        for (uint32_t q : {a, b}) {
          for (size_t j : pending[q]) {
            owner[j] = (int64_t)i;
            pre[i].push_back(j);
          }
          pending[q].clear();
        }
        std::sort(pre[i].begin(), pre[i].end());
        bool open[2] = {true, true};
        const uint32_t qs[2] = {a, b};
        for (size_t j = i + 1; j < N && (open[0] || open[1]); j++) {
          if (owner[j] >= 0) continue;
          const LGate& h = in[j];
          for (int side = 0; side < 2; side++) {
            if (!open[side] || !touches(h, qs[side])) continue;
            if (h.dense_k == 1 && h.dense_q[0] == qs[side]) {
              owner[j] = (int64_t)i;
              post[i].push_back(j);
            } else {
              open[side] = false;
            }
          }
        }
        continue;
      }
    }
    for (uint32_t q : g.nd) pending[q].clear();
    for (uint32_t q : g.dg) pending[q].clear();
  }
  std::vector<LGate> out;
  for (size_t i = 0; i < N; i++) {
    if (owner[i] >= 0) continue;
    if (!root[i]) {
      out.push_back(in[i]);
      continue;
    }
    const LGate& g = in[i];
    const uint32_t a = g.dense_q[0], b = g.dense_q[1];
    double M[32], L[32];
    std::memset(M, 0, sizeof(M));
    for (int d = 0; d < 4; d++) M[2 * (d * 4 + d)] = 1.0;
    for (size_t j : pre[i]) {
      lift2(in[j].dense, in[j].dense_q[0] == a, L);
      mul4(L, M, M);
    }
    mul4(g.dense, M, M);
    for (size_t j : post[i]) {
      lift2(in[j].dense, in[j].dense_q[0] == a, L);
      mul4(L, M, M);
    }
    LGate f;
    f.kind = LGate::G2Q;
    f.qa = a;
    f.qb = b;
    f.nd = {a, b};
    std::memcpy(f.m, M, sizeof(M));
    f.cost = 2.0;
    out.push_back(std::move(f));
  }
  return out;
}

struct PlannerCtx {
  uint32_t n, m, L;
  uint32_t G;        // register bits per thread (amplitudes per thread = 2^G)
  uint32_t feat;     // extra kFeat* bits for every pass (non-temporal streaming)
  double budget;
};

// Greedy dependency-respecting selection: a gate can join the current group
// if none of its qubits is blocked by an earlier gate left out of the group.
// Gates left out block their non-diagonal qubits for everything later and
// their diagonal qubits for later non-diagonal use (diagonal actions on a
// shared qubit commute).
struct Blocker {
  uint64_t any = 0, nd = 0;
  bool admits(uint64_t ndm, uint64_t dgm) const { return !(ndm & (any | nd)) && !(dgm & any); }
  void block(uint64_t ndm, uint64_t dgm) {
    any |= ndm;
    nd |= dgm;
  }
};

// PhaseSlots accumulator of a term whose register part is `rm` (G bits).
uint32_t phase_class(uint32_t rm, uint32_t G, bool* general) {
  *general = false;
  const int c = popc(rm);
  if (c == 0) return 0;
  const uint32_t g = (uint32_t)__builtin_ctz(rm);
  if (c == 1) return 1 + g;
  if (c == 2) {
    const uint32_t h = 31u - (uint32_t)__builtin_clz(rm);
    return 1 + G + g * G - g * (g + 1) / 2 + (h - g - 1);
  }
  *general = true;
  return 0;
}

// FP32 plans keep op matrices as floats (the first `count` slots of the
// matrix area), so the kernel reads them without conversion.
void narrow_matrix(DevOp& op, int count, uint32_t precision) {
  if (precision != ALEPH_HIP_F32) return;
  float f[32];
  for (int i = 0; i < count; i++) f[i] = (float)op.m[i];
  std::memset(op.m, 0, sizeof(op.m));
  std::memcpy(op.m, f, (size_t)count * sizeof(float));
}

void swap_2q_matrix(double* m) {
  // index (a b) -> (b a): swap bit 0 and bit 1 of rows and columns
  double t[32];
  auto sw = [](int x) { return ((x & 1) << 1) | ((x >> 1) & 1); };
  for (int r = 0; r < 4; r++)
    for (int c = 0; c < 4; c++) {
      t[2 * (r * 4 + c)] = m[2 * (sw(r) * 4 + sw(c))];
      t[2 * (r * 4 + c) + 1] = m[2 * (sw(r) * 4 + sw(c)) + 1];
    }
  std::memcpy(m, t, sizeof(t));
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

template <typename T>
size_t push_matrix(std::vector<unsigned char>& payload, const std::vector<double>& mat) {
  const size_t off = align256(payload.size());
  payload.resize(off + mat.size() * sizeof(T));
  T* dst = reinterpret_cast<T*>(payload.data() + off);
  for (size_t i = 0; i < mat.size(); i++) dst[i] = (T)mat[i];
  return off;
}

template <typename V>
size_t push_raw(std::vector<unsigned char>& payload, const std::vector<V>& v) {
  const size_t off = align256(payload.size());
  payload.resize(off + v.size() * sizeof(V) + 8);
  if (!v.empty()) std::memcpy(payload.data() + off, v.data(), v.size() * sizeof(V));
  return off;
}

void add_standalone(Plan& p, const LGate& g) {
  Step s;
  if (g.kind == LGate::KQ) {
    s.kind = Step::KQ;
    s.k = g.k;
    for (uint32_t j = 0; j < g.k; j++) s.qubits[j] = g.kq_qubits[j];
    s.cmask = qmask(g.dg) & ~qmask(g.nd);
    s.dev_off = p.precision == ALEPH_HIP_F64 ? push_matrix<double>(p.payload, g.kq_mat)
                                             : push_matrix<float>(p.payload, g.kq_mat);
  } else {
    s.kind = Step::POLY;
    s.count = (uint32_t)g.poly_angles.size();
    s.dev_off = push_raw(p.payload, g.poly_angles);
    s.dev_off2 = push_raw(p.payload, g.poly_conds);
    s.dev_off3 = push_raw(p.payload, g.poly_off);
  }
  p.steps.push_back(s);
}

// Build one fused pass from the chosen gates.
void emit_fused_pass(Plan& p, const PlannerCtx& pc, uint64_t S, const std::vector<const LGate*>& gates) {
  // chunk qubits: sorted physical positions
  std::vector<uint32_t> pos;
  for (uint32_t q = 0; q < pc.n; q++)
    if ((S >> q) & 1ull) pos.push_back(q);
  uint32_t local_of[64];
  for (uint32_t i = 0; i < 64; i++) local_of[i] = UINT32_MAX;
  for (uint32_t i = 0; i < pos.size(); i++) local_of[pos[i]] = i;
  auto local_mask = [&](uint64_t phys) {
    uint32_t lm = 0;
    for (uint32_t q = 0; q < 64; q++)
      if ((phys >> q) & 1ull) lm |= 1u << local_of[q];
    return lm;
  };
  const uint32_t m = (uint32_t)pos.size();
  const uint32_t low_local = (1u << pc.L) - 1u;  // local bits kept as thread bits at HBM stages

  // ---- stage formation ----
  std::vector<std::pair<uint32_t, std::vector<const LGate*>>> stage_list;  // (G local mask, ops)
  std::vector<const LGate*> rem = gates;
  bool first = true;
  while (!rem.empty()) {
    auto form = [&](bool avoid_low) {
      Blocker bl;
      uint32_t G = 0;
      std::vector<const LGate*> take, left;
      for (const LGate* g : rem) {
        const uint64_t ndm = qmask(g->nd), dgm = qmask(g->dg);
        const uint32_t ndl = local_mask(ndm);
        bool ok = bl.admits(ndm, dgm) && popc(G | ndl) <= (int)pc.G;
        if (ok && avoid_low && (ndl & low_local)) ok = false;
        if (ok) {
          take.push_back(g);
          G |= ndl;
        } else {
          bl.block(ndm, dgm);
          left.push_back(g);
        }
      }
      return std::make_tuple(G, take, left);
    };
    auto [G, take, left] = form(first && pc.L > 0);
    bool has_dense = false;
    for (const LGate* g : take) has_dense |= !g->nd.empty();
    if (first && pc.L > 0 && !has_dense) std::tie(G, take, left) = form(false);
    // pad the register set to G local bits, preferring high local bits
    for (int b = (int)m - 1; b >= 0 && popc(G) < (int)pc.G; b--)
      if (!(G & (1u << b))) G |= 1u << b;
    stage_list.emplace_back(G, take);
    rem = left;
    first = false;
  }
  if (stage_list.empty()) stage_list.emplace_back(0u, std::vector<const LGate*>{});
  // Re-layout stages so that HBM stages keep the low chunk bits as thread bits.
  const uint32_t top = ((1u << m) - 1u) & ~((1u << (m - pc.G)) - 1u);
  if (pc.L > 0 && (stage_list.front().first & low_local))
    stage_list.insert(stage_list.begin(), {top, {}});
  if (pc.L > 0 && (stage_list.back().first & low_local)) stage_list.push_back({top, {}});

  PassDev pd{};
  pd.m = m;
  pd.n_stages = (uint32_t)stage_list.size();
  pd.stage_begin = (uint32_t)p.stages.size();
  for (uint32_t i = 0; i < m; i++) pd.local_pos[i] = pos[i];
  pd.reg_bits = pc.G;
  pd.features = pc.feat;

  for (auto& [G, sgates] : stage_list) {
    StageDev sd{};
    uint32_t reg_local[kMaxRegBits];
    int rcount = 0;
    for (uint32_t b = 0; b < m; b++)
      if (G & (1u << b)) reg_local[rcount++] = b;
    uint32_t reg_of_local[32];
    for (int i = 0; i < 32; i++) reg_of_local[i] = UINT32_MAX;
    uint64_t rphys = 0;
    for (int g = 0; g < (int)pc.G; g++) {
      reg_of_local[reg_local[g]] = (uint32_t)g;
      sd.rl[g] = 1u << reg_local[g];
      sd.rg[g] = 1ull << pos[reg_local[g]];
      rphys |= sd.rg[g];
    }
    uint32_t j = 0;
    for (uint32_t b = 0; b < m; b++) {
      if (G & (1u << b)) continue;
      sd.tl[j] = 1u << b;
      sd.tg[j] = 1ull << pos[b];
      j++;
    }
    sd.n_thread_bits = j;
    {
      // Same fold as the device's swz<SW>: bits >= SW XORed into the low SW bits.
      const uint32_t SW = p.precision == ALEPH_HIP_F64 ? 3u : 4u;
      for (uint32_t r = 0; r < (1u << pc.G); r++) {
        uint64_t go = 0;
        uint32_t lo = 0;
        for (uint32_t g = 0; g < pc.G; g++)
          if (r & (1u << g)) {
            go |= sd.rg[g];
            lo |= sd.rl[g];
          }
        const uint32_t h = lo >> SW;
        sd.go[r] = go;
        sd.ls[r] = lo ^ ((h ^ (h >> SW) ^ (h >> (2 * SW)) ^ (h >> (3 * SW))) & ((1u << SW) - 1u));
        sd.lb[r] = sd.ls[r] * 8u;
      }
    }
    sd.op_begin = (uint32_t)p.ops.size();
    auto reg_mask = [&](uint64_t phys) {
      uint32_t rm = 0;
      for (int g = 0; g < (int)pc.G; g++)
        if (phys & sd.rg[g]) rm |= 1u << g;
      return rm;
    };
    std::vector<PhaseTerm> pend_reg, pend_gen;
    // A term whose mask has no chunk bit outside the register bits depends
    // only on the chunk number: the same for every thread of a workgroup.
    auto thread_dep = [&](const PhaseTerm& t) { return (t.mask_out & S) != 0; };
    auto flush_phase = [&]() {
      if (pend_reg.empty() && pend_gen.empty()) return;
      DevOp ph{};
      ph.type = OP_PHASE;
      // Sort by accumulator class, and within a class the workgroup-uniform
      // terms first.  The matrix slots hold the class boundaries
      // (bounds[0..kCount]) and then each class's first thread-dependent
      // term (split[0..kCount-1]), all relative to term_off.
      std::stable_sort(pend_reg.begin(), pend_reg.end(), [&](const PhaseTerm& x, const PhaseTerm& y) {
        if (x.cls != y.cls) return x.cls < y.cls;
        return !thread_dep(x) && thread_dep(y);
      });
      const uint32_t n_classes = 1 + pc.G + pc.G * (pc.G - 1) / 2;
      uint32_t* bounds = reinterpret_cast<uint32_t*>(ph.m);
      uint32_t* split = bounds + n_classes + 1;
      static_assert(sizeof(DevOp::m) >= (2 * (1 + kMaxRegBits + kMaxRegBits * (kMaxRegBits - 1) / 2) + 1) * 4,
                    "phase class tables fit the matrix slots");
      for (uint32_t k = 0, t = 0; k <= n_classes; k++) {
        while (t < pend_reg.size() && pend_reg[t].cls < k) t++;
        bounds[k] = t;
      }
      for (uint32_t k = 0; k < n_classes; k++) {
        uint32_t t = bounds[k];
        while (t < bounds[k + 1] && !thread_dep(pend_reg[t])) t++;
        split[k] = t;
      }
      ph.term_off = (uint32_t)p.terms.size();
      ph.n_terms = (uint32_t)pend_reg.size();
      ph.b = (uint32_t)pend_gen.size();
      p.terms.insert(p.terms.end(), pend_reg.begin(), pend_reg.end());
      p.terms.insert(p.terms.end(), pend_gen.begin(), pend_gen.end());
      p.ops.push_back(ph);
      pend_reg.clear();
      pend_gen.clear();
    };
    for (const LGate* g : sgates) {
      DevOp op{};
      if (g->kind != LGate::TERMS) flush_phase();
      switch (g->kind) {
        case LGate::G1Q:
          op.type = OP_1Q;
          op.a = reg_of_local[local_of[g->qa]];
          op.cmask = g->cmask;
          std::memcpy(op.m, g->m, 8 * sizeof(double));
          narrow_matrix(op, 8, p.precision);
          p.ops.push_back(op);
          break;
        case LGate::G2Q: {
          op.type = OP_2Q;
          uint32_t ra = reg_of_local[local_of[g->qa]], rb = reg_of_local[local_of[g->qb]];
          std::memcpy(op.m, g->m, 32 * sizeof(double));
          if (ra < rb) {
            std::swap(ra, rb);
            swap_2q_matrix(op.m);
          }
          op.a = ra;
          op.b = rb;
          op.cmask = g->cmask;
          narrow_matrix(op, 32, p.precision);
          p.ops.push_back(op);
          break;
        }
        case LGate::SIGN:
          op.type = OP_SIGN;
          op.cmask = g->sign_mask & ~rphys;
          op.a = reg_mask(g->sign_mask);
          p.ops.push_back(op);
          break;
        case LGate::DIAGX:
          op.type = OP_DIAG;
          op.ma = 1ull << g->qa;
          op.mb = g->n_diag_targets == 2 ? (1ull << g->qb) : 0;
          op.cmask = g->cmask;
          std::memcpy(op.m, g->m, 8 * sizeof(double));
          narrow_matrix(op, 8, p.precision);
          p.ops.push_back(op);
          break;
        case LGate::TERMS:
          for (const auto& [mask, angle] : g->terms) {
            PhaseTerm t{};
            t.mask_out = mask & ~rphys;
            t.rmask = reg_mask(mask);
            t.angle = angle;
            t.c = std::cos(angle);
            t.s = std::sin(angle);
            bool general = false;
            t.cls = phase_class(t.rmask, pc.G, &general);
            (general ? pend_gen : pend_reg).push_back(t);
          }
          break;
        default:
          break;
      }
    }
    flush_phase();
    sd.op_end = (uint32_t)p.ops.size();
    for (uint32_t o = sd.op_begin; o < sd.op_end; o++) {
      if (p.ops[o].type == OP_2Q) pd.features |= kFeat2Q;
      if (p.ops[o].type == OP_PHASE) pd.features |= kFeatPhase;
    }
    p.stages.push_back(sd);
  }
  Step s;
  s.kind = Step::FUSED;
  s.pass = (uint32_t)p.passes.size();
  s.m = m;
  s.n_stages = pd.n_stages;
  p.passes.push_back(pd);
  p.steps.push_back(s);
}

void plan_rule(Plan& p, std::vector<LGate> lg, uint32_t n, const PlanOptions& o) {
  PlannerCtx pc;
  pc.n = n;
  // 2^12-amplitude chunks (256-thread workgroups) for both precisions: FP32
  // chunks of 2^13 need 512-thread workgroups, of which one fits per CU,
  // and ran brickwall passes 1.8x slower on MI355X.
  const uint32_t mdef = 12u;
  pc.m = std::min<uint32_t>(o.local_qubits ? o.local_qubits : mdef, n);
  pc.m = std::min<uint32_t>(pc.m, kMaxLocalBits);
  pc.G = o.reg_bits ? o.reg_bits : (uint32_t)kRegBits;
  pc.feat = o.nontemporal ? kFeatNontemporal : 0u;
  // Four low qubits in every chunk: runs of 16 consecutive amplitudes (256 B
  // FP64, 128 B FP32).  With three, a pass whose other chunk qubits are all
  // high (20..28 at n=30) streamed at 9.1 ms (FP64) and 7.6 ms (FP32) against
  // 6.8 / 3.3 ms for the skeleton; at n=30 the QFT ran 45.8 -> 43.5 ms and
  // an n=28 HEA 32.8 -> 26.7 ms, the brickwall unchanged.
  pc.L = std::min<uint32_t>(o.coalesce_bits ? o.coalesce_bits : 4u, pc.m - pc.G);
  pc.budget = o.max_ops ? (double)o.max_ops : 24.0;
  const uint64_t all = n >= 64 ? ~0ull : ((1ull << n) - 1ull);
  std::vector<char> done(lg.size(), 0);
  size_t first = 0;
  while (true) {
    while (first < lg.size() && done[first]) first++;
    if (first == lg.size()) break;
    if (lg[first].standalone()) {
      add_standalone(p, lg[first]);
      done[first] = 1;
      continue;
    }
    uint64_t S = (1ull << pc.L) - 1ull;
    Blocker bl;
    double cost = 0.0;
    std::vector<const LGate*> take;
    for (size_t i = first; i < lg.size(); i++) {
      if (done[i]) continue;
      const LGate& g = lg[i];
      const uint64_t ndm = qmask(g.nd), dgm = qmask(g.dg);
      bool ok = !g.standalone() && bl.admits(ndm, dgm) && popc(S | ndm) <= (int)pc.m;
      if (ok && !take.empty() && cost + g.cost > pc.budget) {
        // budget spent: close the pass here
        break;
      }
      if (ok) {
        take.push_back(&g);
        done[i] = 1;
        S |= ndm;
        cost += g.cost;
      } else {
        bl.block(ndm, dgm);
        if ((bl.any & all) == all) break;
      }
    }
    // pad the chunk with the lowest unused qubits (longer contiguous runs)
    for (uint32_t q = 0; q < n && popc(S) < (int)pc.m; q++) S |= 1ull << q;
    emit_fused_pass(p, pc, S, take);
  }
}

void plan_unfused(Plan& p, const std::vector<ProgInstr>& prog) {
  for (const ProgInstr& pi : prog) {
    if (pi.type == ProgInstr::BARRIER) continue;
    Step s;
    if (pi.type == ProgInstr::DIAG_PHASE) {
      std::vector<double> ang = pi.angles;
      std::vector<uint64_t> conds;
      std::vector<uint32_t> off{0};
      for (const auto& cs : pi.conds) {
        for (uint64_t c : cs) conds.push_back(c);
        off.push_back((uint32_t)conds.size());
      }
      s.kind = Step::POLY;
      s.count = (uint32_t)ang.size();
      s.dev_off = push_raw(p.payload, ang);
      s.dev_off2 = push_raw(p.payload, conds);
      s.dev_off3 = push_raw(p.payload, off);
      p.steps.push_back(s);
      continue;
    }
    const DenseGate& d = pi.gate;
    const uint32_t k = (uint32_t)d.qubits.size();
    s.cmask = qmask(d.controls);
    s.k = k;
    for (uint32_t j = 0; j < k && j < 8; j++) s.qubits[j] = d.qubits[j];
    if (d.diagonal) {
      std::vector<double> diag(2 * (size_t)d.dim);
      for (uint32_t j = 0; j < d.dim; j++) {
        diag[2 * j] = d.m[(size_t)j * d.dim + j].real();
        diag[2 * j + 1] = d.m[(size_t)j * d.dim + j].imag();
      }
      s.kind = Step::DIAG;
      s.dev_off = push_raw(p.payload, diag);
    } else if (d.dim == 2 || d.dim == 4) {
      s.kind = d.dim == 2 ? Step::G1Q : Step::G2Q;
      for (uint32_t e = 0; e < d.dim * d.dim; e++) {
        s.mat8[2 * e] = d.m[e].real();
        s.mat8[2 * e + 1] = d.m[e].imag();
      }
    } else {
      std::vector<double> mat(2 * (size_t)d.dim * d.dim);
      for (size_t e = 0; e < (size_t)d.dim * d.dim; e++) {
        mat[2 * e] = d.m[e].real();
        mat[2 * e + 1] = d.m[e].imag();
      }
      s.kind = Step::KQ;
      s.dev_off = p.precision == ALEPH_HIP_F64 ? push_matrix<double>(p.payload, mat)
                                               : push_matrix<float>(p.payload, mat);
    }
    p.steps.push_back(s);
  }
}

}  // namespace

std::vector<ProgInstr> fuse_diagonal_runs(const std::vector<ProgInstr>& prog, uint32_t n) {
  // Runs of {diagonal gate, bare CNOT, DiagonalPhase} that contain a CNOT
  // and whose CNOTs compose to the identity become one DiagonalPhase: each
  // physical qubit carries the parity mask of input bits it holds, and a
  // diagonal gate's Moebius terms become parity terms over those masks
  // (the run semantics of crates/aleph-ir/src/passes/fuse_diagonal.rs:1-80).
  auto is_cnot = [](const ProgInstr& pi) {
    return pi.type == ProgInstr::GATE && pi.gate.kind == ALEPH_HIP_GATE_CNOT && pi.gate.controls.empty();
  };
  auto member = [&](const ProgInstr& pi) {
    return pi.type == ProgInstr::DIAG_PHASE || (pi.type == ProgInstr::GATE && (pi.gate.diagonal || is_cnot(pi)));
  };
  std::vector<ProgInstr> out;
  size_t i = 0;
  while (i < prog.size()) {
    if (!member(prog[i])) {
      out.push_back(prog[i++]);
      continue;
    }
    size_t j = i;
    bool has_cnot = false;
    while (j < prog.size() && member(prog[j])) has_cnot |= is_cnot(prog[j++]);
    std::vector<uint64_t> lin(n);
    for (uint32_t q = 0; q < n; q++) lin[q] = 1ull << q;
    auto lmask = [&](uint64_t phys) {
      uint64_t v = 0;
      for (uint64_t r = phys; r; r &= r - 1) v ^= lin[__builtin_ctzll(r)];
      return v;
    };
    ProgInstr fused;
    fused.type = ProgInstr::DIAG_PHASE;
    for (size_t t = i; has_cnot && t < j; t++) {
      const ProgInstr& pi = prog[t];
      if (is_cnot(pi)) {
        lin[pi.gate.qubits[1]] ^= lin[pi.gate.qubits[0]];
        continue;
      }
      if (pi.type == ProgInstr::DIAG_PHASE) {
        for (size_t u = 0; u < pi.angles.size(); u++) {
          std::vector<uint64_t> cs;
          for (uint64_t c : pi.conds[u]) cs.push_back(lmask(c));
          fused.conds.push_back(std::move(cs));
          fused.angles.push_back(pi.angles[u]);
        }
        continue;
      }
      const DenseGate& d = pi.gate;
      std::vector<std::pair<uint64_t, double>> terms;
      diag_to_terms(d, qmask(d.controls), terms);
      // Constant phase of the targets' |0..0> entry: diag_to_terms drops
      // a zero beta only, so the S = 0 term carries it when nonzero.
      for (const auto& tm : terms) {
        std::vector<uint64_t> cs;
        for (uint64_t r = tm.first; r; r &= r - 1) cs.push_back(lin[__builtin_ctzll(r)]);
        fused.conds.push_back(std::move(cs));
        fused.angles.push_back(tm.second);
      }
    }
    bool identity = has_cnot;
    for (uint32_t q = 0; identity && q < n; q++) identity = lin[q] == (1ull << q);
    if (identity) {
      out.push_back(std::move(fused));
    } else {
      for (size_t t = i; t < j; t++) out.push_back(prog[t]);
    }
    i = j;
  }
  return out;
}

namespace {

// Apply gate `g` (targets MSB first, optional controls) to a vector over the
// block's qubits `bq` (ascending; local bit j is qubit bq[j]).
void apply_to_local(std::vector<cplx>& v, const std::vector<uint32_t>& bq, const DenseGate& g) {
  auto local = [&](uint32_t q) {
    return (uint32_t)(std::lower_bound(bq.begin(), bq.end(), q) - bq.begin());
  };
  const uint32_t kg = (uint32_t)g.qubits.size();
  uint32_t tmask = 0, cmask = 0;
  std::vector<uint32_t> off(g.dim);
  for (uint32_t t = 0; t < g.dim; t++) {
    uint32_t o = 0;
    for (uint32_t j = 0; j < kg; j++)
      if ((t >> j) & 1u) o |= 1u << local(g.qubits[kg - 1 - j]);
    off[t] = o;
  }
  for (uint32_t q : g.qubits) tmask |= 1u << local(q);
  for (uint32_t q : g.controls) cmask |= 1u << local(q);
  std::vector<cplx> in(g.dim);
  const uint32_t D = (uint32_t)v.size();
  for (uint32_t x = 0; x < D; x++) {
    if ((x & tmask) != 0 || (x & cmask) != cmask) continue;
    for (uint32_t t = 0; t < g.dim; t++) in[t] = v[x | off[t]];
    for (uint32_t r = 0; r < g.dim; r++) {
      cplx acc = 0;
      for (uint32_t c = 0; c < g.dim; c++) acc += g.m[(size_t)r * g.dim + c] * in[c];
      v[x | off[r]] = acc;
    }
  }
}

struct OpenBlock {
  std::vector<uint32_t> qubits;            // ascending
  std::vector<DenseGate> gates;            // in order
};

DenseGate block_gate(const OpenBlock& b) {
  const uint32_t k = (uint32_t)b.qubits.size();
  const uint32_t D = 1u << k;
  DenseGate out;
  out.kind = k == 1 ? ALEPH_HIP_GATE_UNITARY1Q : k == 2 ? ALEPH_HIP_GATE_UNITARY2Q : ALEPH_HIP_GATE_UNITARYKQ;
  out.qubits.assign(b.qubits.rbegin(), b.qubits.rend());   // MSB first
  out.dim = D;
  out.m.assign((size_t)D * D, cplx(0, 0));
  std::vector<cplx> col(D);
  for (uint32_t c = 0; c < D; c++) {
    std::fill(col.begin(), col.end(), cplx(0, 0));
    col[c] = 1;
    for (const DenseGate& g : b.gates) apply_to_local(col, b.qubits, g);
    for (uint32_t r = 0; r < D; r++) out.m[(size_t)r * D + c] = col[r];
  }
  out.diagonal = matrix_is_diagonal(out.m, D);
  return out;
}

}  // namespace

std::vector<ProgInstr> fuse_dense_blocks(const std::vector<ProgInstr>& prog, uint32_t width) {
  // Greedy dense blocking: a gate joins the open blocks it touches when
  // their qubits together with its own number at most `width`; otherwise
  // those blocks are emitted first.  Open blocks are disjoint, so they
  // commute and may be emitted in any order.  Blocks of one gate are
  // emitted unchanged.  (The reference's rule fuser, FuseKq,
  // crates/aleph-ir/src/passes/fuse_kq.rs:185-390, grows blocks over a
  // dependency window with its own emit rule; this is a simpler greedy.)
  std::vector<ProgInstr> out;
  std::vector<OpenBlock> open;
  auto emit = [&](size_t i) {
    ProgInstr pi;
    pi.type = ProgInstr::GATE;
    pi.gate = open[i].gates.size() == 1 ? open[i].gates[0] : block_gate(open[i]);
    out.push_back(std::move(pi));
    open.erase(open.begin() + (long)i);
  };
  auto flush_all = [&] {
    while (!open.empty()) emit(0);
  };
  for (const ProgInstr& pi : prog) {
    if (pi.type != ProgInstr::GATE) {
      flush_all();
      out.push_back(pi);
      continue;
    }
    const DenseGate& g = pi.gate;
    std::vector<uint32_t> q = g.qubits;
    q.insert(q.end(), g.controls.begin(), g.controls.end());
    std::sort(q.begin(), q.end());
    std::vector<size_t> touch;
    std::vector<uint32_t> uni = q;
    for (size_t i = 0; i < open.size(); i++) {
      bool hit = false;
      for (uint32_t x : open[i].qubits)
        if (std::binary_search(q.begin(), q.end(), x)) hit = true;
      if (hit) {
        touch.push_back(i);
        uni.insert(uni.end(), open[i].qubits.begin(), open[i].qubits.end());
      }
    }
    std::sort(uni.begin(), uni.end());
    uni.erase(std::unique(uni.begin(), uni.end()), uni.end());
    if (q.size() > width) {
      for (size_t t = touch.size(); t-- > 0;) emit(touch[t]);
      out.push_back(pi);
      continue;
    }
    if (uni.size() > width) {
      for (size_t t = touch.size(); t-- > 0;) emit(touch[t]);
      open.push_back(OpenBlock{q, {g}});
      continue;
    }
    OpenBlock merged{uni, {}};
    for (size_t t : touch)
      merged.gates.insert(merged.gates.end(), open[t].gates.begin(), open[t].gates.end());
    merged.gates.push_back(g);
    for (size_t t = touch.size(); t-- > 0;) open.erase(open.begin() + (long)touch[t]);
    open.push_back(std::move(merged));
  }
  flush_all();
  return out;
}

namespace {
// A run of uncontrolled SWAPs (and barriers) that ends the program is a
// qubit permutation sigma (sigma[q]: the position bit q's value ends at).
// When it is an involution with at least two SWAPs (disjoint transpositions,
// the QFT's final reversal among them) it runs as one in-place sweep instead
// of swap passes: returns where the run starts, else prog.size().
size_t trailing_swap_involution(const std::vector<ProgInstr>& prog, uint32_t n, std::vector<uint32_t>& sigma) {
  size_t cut = prog.size(), swaps = 0;
  while (cut > 0) {
    const ProgInstr& pi = prog[cut - 1];
    const bool swap = pi.type == ProgInstr::GATE && pi.gate.kind == ALEPH_HIP_GATE_SWAP && pi.gate.controls.empty();
    if (pi.type != ProgInstr::BARRIER && !swap) break;
    swaps += swap ? 1 : 0;
    cut--;
  }
  sigma.resize(n);
  for (uint32_t q = 0; q < n; q++) sigma[q] = q;
  for (size_t i = cut; i < prog.size(); i++) {
    if (prog[i].type != ProgInstr::GATE) continue;
    const uint32_t a = prog[i].gate.qubits[0], b = prog[i].gate.qubits[1];
    for (uint32_t& v : sigma) v = v == a ? b : (v == b ? a : v);
  }
  bool identity = true, involution = true;
  for (uint32_t q = 0; q < n; q++) {
    identity &= sigma[q] == q;
    involution &= sigma[sigma[q]] == q;
  }
  return swaps >= 2 && involution && !identity ? cut : prog.size();
}

// The permutation kernel holds tiles over the low min(5, n) qubits and their
// images under sigma in LDS; uncontrolled one-qubit gates on those qubits
// that directly precede the permutation are applied there, in the same
// sweep (the QFT's last Hadamard lands on such a qubit).
size_t tile_gates_before(const std::vector<ProgInstr>& prog, size_t cut, uint32_t n, const std::vector<uint32_t>& sigma) {
  uint64_t tile = 0;
  for (uint32_t q = 0; q < std::min<uint32_t>(5u, n); q++) tile |= (1ull << q) | (1ull << sigma[q]);
  size_t start = cut, gates = 0;
  while (start > 0 && gates < kPermMaxGates) {
    const ProgInstr& pi = prog[start - 1];
    if (pi.type == ProgInstr::BARRIER) {
      start--;
      continue;
    }
    if (pi.type != ProgInstr::GATE || pi.gate.qubits.size() != 1 || !pi.gate.controls.empty() ||
        pi.gate.dim != 2 || !((tile >> pi.gate.qubits[0]) & 1ull))
      break;
    start--;
    gates++;
  }
  return start;
}

// Whether gate b undoes gate a exactly: the same qubits in the same order,
// the same controls, and b's matrix equal to a's conjugate transpose entry
// for entry (Gate::inverse's closed forms give that for the self-inverse
// kinds, S/Sdg, T/Tdg, Iswap/IswapDg and negated rotation angles).
bool exact_inverse(const DenseGate& a, const DenseGate& b) {
  if (a.qubits != b.qubits || a.dim != b.dim || a.m.size() != b.m.size()) return false;
  std::vector<uint32_t> ca = a.controls, cb = b.controls;
  std::sort(ca.begin(), ca.end());
  std::sort(cb.begin(), cb.end());
  if (ca != cb) return false;
  const uint32_t d = a.dim;
  for (uint32_t r = 0; r < d; r++)
    for (uint32_t c = 0; c < d; c++)
      if (b.m[(size_t)r * d + c] != std::conj(a.m[(size_t)c * d + r])) return false;
  return true;
}

// CancelInversePairs (crates/aleph-ir/src/passes/cancel.rs:1-80) on the
// decoded program: one forward pass with a stack of live instruction
// indices per qubit.  A gate cancels with the live instruction on top of
// every stack it uses when that instruction is a gate it undoes exactly;
// both go.  Diagonal phases and barriers never cancel, so nothing cancels
// across them (a barrier blocks every qubit: the decoded form keeps no
// qubit list).
std::vector<ProgInstr> cancel_inverse_pairs(const std::vector<ProgInstr>& prog, uint32_t n, uint32_t* cancelled) {
  std::vector<std::vector<uint32_t>> live(n);
  std::vector<char> dead(prog.size(), 0);
  std::vector<uint32_t> sup;
  uint32_t count = 0;
  for (size_t i = 0; i < prog.size(); i++) {
    const ProgInstr& pi = prog[i];
    sup.clear();
    if (pi.type == ProgInstr::GATE) {
      sup = pi.gate.qubits;
      sup.insert(sup.end(), pi.gate.controls.begin(), pi.gate.controls.end());
    } else if (pi.type == ProgInstr::DIAG_PHASE) {
      uint64_t mk = 0;
      for (const auto& c : pi.conds)
        for (uint64_t v : c) mk |= v;
      for (uint32_t q = 0; q < n; q++)
        if ((mk >> q) & 1ull) sup.push_back(q);
    } else {
      for (uint32_t q = 0; q < n; q++) sup.push_back(q);
    }
    if (pi.type == ProgInstr::GATE && !sup.empty()) {
      int64_t cand = -1;
      bool shared = true;
      for (uint32_t q : sup) {
        if (live[q].empty()) {
          shared = false;
          break;
        }
        const int64_t top = live[q].back();
        if (cand < 0) cand = top;
        else if (cand != top) {
          shared = false;
          break;
        }
      }
      if (shared && cand >= 0 && prog[(size_t)cand].type == ProgInstr::GATE &&
          exact_inverse(prog[(size_t)cand].gate, pi.gate)) {
        dead[(size_t)cand] = 1;
        dead[i] = 1;
        for (uint32_t q : sup) live[q].pop_back();
        count += 2;
        continue;
      }
    }
    for (uint32_t q : sup) live[q].push_back((uint32_t)i);
  }
  if (cancelled) *cancelled = count;
  if (!count) return prog;
  std::vector<ProgInstr> out;
  out.reserve(prog.size() - count);
  for (size_t i = 0; i < prog.size(); i++)
    if (!dead[i]) out.push_back(prog[i]);
  return out;
}
}  // namespace

std::unique_ptr<Plan> build_plan(const std::vector<ProgInstr>& prog_in, uint32_t n, uint32_t precision,
                                 const PlanOptions& opts) {
  const auto t0 = std::chrono::steady_clock::now();
  auto p = std::make_unique<Plan>();
  p->n = n;
  p->precision = precision;
  p->opts = opts;
  uint32_t gates = 0;
  for (const ProgInstr& pi : prog_in) gates += pi.type == ProgInstr::GATE ? 1 : 0;
  const uint32_t G = opts.reg_bits ? opts.reg_bits : (uint32_t)kRegBits;
  if (opts.planner == ALEPH_HIP_PLANNER_UNFUSED || n < G) {
    plan_unfused(*p, prog_in);
  } else {
    // Only states wider than a chunk: inside one chunk the swaps are register renames of the last pass.
    std::vector<uint32_t> sigma;
    const uint32_t m_chunk = std::min<uint32_t>(opts.local_qubits ? opts.local_qubits : 12u, (uint32_t)kMaxLocalBits);
    const size_t cut = n > m_chunk ? trailing_swap_involution(prog_in, n, sigma) : prog_in.size();
    const size_t body = cut < prog_in.size() ? tile_gates_before(prog_in, cut, n, sigma) : cut;
    const std::vector<ProgInstr> prog =
        cancel_inverse_pairs(std::vector<ProgInstr>(prog_in.begin(), prog_in.begin() + (ptrdiff_t)body), n,
                             &p->stats.cancelled_gates);
    std::vector<ProgInstr> pre = opts.fuse_diag ? fuse_diagonal_runs(prog, n) : prog;
    if (opts.kq_width >= 3) pre = fuse_dense_blocks(pre, std::min<uint32_t>(opts.kq_width, kMaxFusedK));
    std::vector<LGate> lg = lower_program(pre, n);
    if (opts.fuse_1q) lg = fuse_1q_runs(lg);
    if (opts.fuse_2q) lg = fuse_pairs(lg);
    plan_rule(*p, std::move(lg), n, opts);
    if (cut < prog_in.size()) {
      Step st;
      st.kind = Step::BITPERM;
      st.count = n;
      st.dev_off = push_raw(p->payload, sigma);
      std::vector<uint32_t> gq;
      std::vector<double> gm;
      for (size_t i = body; i < cut; i++) {
        if (prog_in[i].type != ProgInstr::GATE) continue;
        gq.push_back(prog_in[i].gate.qubits[0]);
        for (const cplx& z : prog_in[i].gate.m) {
          gm.push_back(z.real());
          gm.push_back(z.imag());
        }
      }
      st.k = (uint32_t)gq.size();
      st.dev_off2 = push_raw(p->payload, gq.empty() ? std::vector<uint32_t>{0} : gq);
      st.dev_off3 = push_raw(p->payload, gm.empty() ? std::vector<double>{0.0} : gm);
      p->steps.push_back(st);
    }
  }
  const auto t1 = std::chrono::steady_clock::now();
  aleph_hip_plan_stats& st = p->stats;
  st.plan_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  st.input_gates = gates;
  st.n_passes = (uint32_t)p->steps.size();
  st.n_stages = (uint32_t)p->stages.size();
  st.n_ops = (uint32_t)p->ops.size();
  st.n_launches = (uint32_t)p->steps.size();
  const double bpa = precision == ALEPH_HIP_F64 ? 16.0 : 8.0;
  st.sweep_bytes = (double)st.n_passes * 2.0 * std::ldexp(bpa, (int)n);
  return p;
}

void plan_features(const Plan& p, std::vector<float>& rows, std::vector<uint32_t>& launch) {
  for (size_t i = 0; i < p.steps.size(); i++) {
    const Step& s = p.steps[i];
    if (s.kind != Step::FUSED) continue;
    const PassDev& pd = p.passes[s.pass];
    float n1q = 0, n2q = 0, ndiag = 0, nsign = 0, nphase = 0, nreg = 0, ngen = 0, nctl = 0;
    for (uint32_t st = pd.stage_begin; st < pd.stage_begin + pd.n_stages; st++) {
      for (uint32_t o = p.stages[st].op_begin; o < p.stages[st].op_end; o++) {
        const DevOp& op = p.ops[o];
        switch (op.type) {
          case OP_1Q: n1q += 1; break;
          case OP_2Q: n2q += 1; break;
          case OP_DIAG: ndiag += 1; break;
          case OP_SIGN: nsign += 1; break;
          case OP_PHASE:
            nphase += 1;
            nreg += (float)op.n_terms;
            ngen += (float)op.b;
            break;
          default: break;
        }
        if (op.type != OP_SIGN && op.type != OP_PHASE && op.cmask != 0) nctl += 1;
      }
    }
    // The model predicts a launch on 2^kCostRefQubits amplitudes and the
    // planner scales by the state size.  Slot 14 still carries the state's
    // size, log2(bytes) relative to 256 MiB (MI355X's last-level Infinity
    // Cache): a state that stays cache-resident between passes streams
    // faster, so a large pass budget that pays at n = 24 can be
    // compute-bound at n = 30.  Slot 0 is the code variant (1: kernel
    // generated for the pass).  The register width (reg_bits) is not a
    // feature: every learned candidate uses the default of 4.
    const float log2_bytes = (float)p.n + (p.precision == ALEPH_HIP_F64 ? 4.f : 3.f);
    const float f[kPlanFeatures] = {p.opts.jit ? 1.f : 0.f,
                                    p.precision == ALEPH_HIP_F64 ? 1.f : 0.f,
                                    (float)pd.m / 13.f,
                                    (float)pd.n_stages / 8.f,
                                    n1q / 32.f,
                                    n2q / 32.f,
                                    ndiag / 16.f,
                                    nsign / 32.f,
                                    nphase / 8.f,
                                    nreg / 128.f,
                                    ngen / 64.f,
                                    (pd.features & kFeat2Q) ? 1.f : 0.f,
                                    (pd.features & kFeatPhase) ? 1.f : 0.f,
                                    (pd.features & kFeatNontemporal) ? 1.f : 0.f,
                                    (log2_bytes - 28.f) / 6.f,
                                    nctl / 32.f};
    rows.insert(rows.end(), f, f + kPlanFeatures);
    launch.push_back((uint32_t)i);
  }
}

// Standalone steps are priced from their measured rates on MI355X: a dense
// block of k <= 5 qubits or a phase polynomial streams at about 4.8 TB/s,
// an FP64 k = 6 block is matrix-core bound at about twice that time.
double plan_standalone_us(const Plan& p) {
  const double bpa = p.precision == ALEPH_HIP_F64 ? 16.0 : 8.0;
  const double sweep_us = 2.0 * std::ldexp(bpa, (int)p.n) / 4.8e12 * 1e6;
  double us = 0.0;
  for (const Step& s : p.steps) {
    if (s.kind == Step::FUSED) continue;
    us += (s.kind == Step::KQ && s.k >= 6 && p.precision == ALEPH_HIP_F64) ? 2.0 * sweep_us : sweep_us;
  }
  return us;
}

Plan::~Plan() {
  if (d_passes) (void)hipFree(d_passes);
  if (d_stages) (void)hipFree(d_stages);
  if (d_ops) (void)hipFree(d_ops);
  if (d_terms) (void)hipFree(d_terms);
  if (d_payload) (void)hipFree(d_payload);
}

namespace {
template <typename V>
void upload_vec(const std::vector<V>& v, V** dst, hipStream_t s) {
  const size_t bytes = std::max<size_t>(v.size(), 1) * sizeof(V);
  ALEPH_HIP_CHECK(hipMalloc(reinterpret_cast<void**>(dst), bytes));
  if (!v.empty()) ALEPH_HIP_CHECK(hipMemcpyAsync(*dst, v.data(), v.size() * sizeof(V), hipMemcpyHostToDevice, s));
}
}  // namespace

void Plan::upload(hipStream_t s) {
  upload_vec(passes, &d_passes, s);
  upload_vec(stages, &d_stages, s);
  upload_vec(ops, &d_ops, s);
  upload_vec(terms, &d_terms, s);
  upload_vec(payload, &d_payload, s);
  ALEPH_HIP_CHECK(hipStreamSynchronize(s));
  uploaded = true;
}

template <typename T>
static void execute_t(const Plan& p, hipStream_t s, void* state, hipEvent_t* ev) {
  for (size_t i = 0; i < p.steps.size(); i++) {
    const Step& st = p.steps[i];
    if (ev) ALEPH_HIP_CHECK(hipEventRecord(ev[i], s));
    switch (st.kind) {
      case Step::FUSED:
        if (!p.jit.empty() && p.jit[st.pass]) {
          launch_jit_pass(s, p, i, state);
          break;
        }
        launch_fused_pass<T>(s, state, p.n, st.m, st.n_stages, p.passes[st.pass].features,
                             p.passes[st.pass].reg_bits, p.d_passes + st.pass, p.d_stages, p.d_ops, p.d_terms);
        break;
      case Step::KQ:
        launch_kq<T>(s, state, p.n, st.k, st.qubits, st.cmask, p.d_payload + st.dev_off);
        break;
      case Step::DIAG:
        launch_diag<T>(s, state, p.n, st.k, st.qubits, st.cmask,
                       reinterpret_cast<const double*>(p.d_payload + st.dev_off));
        break;
      case Step::POLY:
        launch_phase_poly<T>(s, state, p.n, st.count, reinterpret_cast<const double*>(p.d_payload + st.dev_off),
                             reinterpret_cast<const uint64_t*>(p.d_payload + st.dev_off2),
                             reinterpret_cast<const uint32_t*>(p.d_payload + st.dev_off3));
        break;
      case Step::G1Q:
        launch_1q<T>(s, state, p.n, st.qubits[0], st.cmask, st.mat8);
        break;
      case Step::G2Q:
        launch_2q<T>(s, state, p.n, st.qubits[0], st.qubits[1], st.cmask, st.mat8);
        break;
      case Step::SWAPQ:
        launch_swap<T>(s, state, p.n, st.qubits[0], st.qubits[1], st.cmask);
        break;
      case Step::BITPERM:
        launch_bitperm<T>(s, state, p.n, reinterpret_cast<const uint32_t*>(p.payload.data() + st.dev_off), st.k,
                          reinterpret_cast<const uint32_t*>(p.payload.data() + st.dev_off2),
                          reinterpret_cast<const double*>(p.payload.data() + st.dev_off3));
        break;
    }
  }
  if (ev) ALEPH_HIP_CHECK(hipEventRecord(ev[p.steps.size()], s));
}

void Plan::execute(hipStream_t s, void* state) const {
  if (precision == ALEPH_HIP_F64) execute_t<double>(*this, s, state, nullptr);
  else execute_t<float>(*this, s, state, nullptr);
}

void Plan::execute_with_events(hipStream_t s, void* state, hipEvent_t* ev) const {
  if (precision == ALEPH_HIP_F64) execute_t<double>(*this, s, state, ev);
  else execute_t<float>(*this, s, state, ev);
}

void apply_dense_gate(hipStream_t s, void* state, uint32_t n, uint32_t precision, const DenseGate& d,
                      void* scratch) {
  const uint64_t cm = qmask(d.controls);
  const uint32_t k = (uint32_t)d.qubits.size();
  const bool f64 = precision == ALEPH_HIP_F64;
  if (d.diagonal) {
    std::vector<double> diag(2 * (size_t)d.dim);
    for (uint32_t j = 0; j < d.dim; j++) {
      diag[2 * j] = d.m[(size_t)j * d.dim + j].real();
      diag[2 * j + 1] = d.m[(size_t)j * d.dim + j].imag();
    }
    ALEPH_HIP_CHECK(hipMemcpyAsync(scratch, diag.data(), diag.size() * sizeof(double), hipMemcpyHostToDevice, s));
    if (f64) launch_diag<double>(s, state, n, k, d.qubits.data(), cm, (const double*)scratch);
    else launch_diag<float>(s, state, n, k, d.qubits.data(), cm, (const double*)scratch);
    ALEPH_HIP_CHECK(hipStreamSynchronize(s));
    return;
  }
  if (d.dim == 2 || d.dim == 4) {
    double m[32];
    for (uint32_t e = 0; e < d.dim * d.dim; e++) {
      m[2 * e] = d.m[e].real();
      m[2 * e + 1] = d.m[e].imag();
    }
    if (d.dim == 2) {
      if (f64) launch_1q<double>(s, state, n, d.qubits[0], cm, m);
      else launch_1q<float>(s, state, n, d.qubits[0], cm, m);
    } else {
      if (f64) launch_2q<double>(s, state, n, d.qubits[0], d.qubits[1], cm, m);
      else launch_2q<float>(s, state, n, d.qubits[0], d.qubits[1], cm, m);
    }
    return;
  }
  std::vector<double> md(2 * (size_t)d.dim * d.dim);
  for (size_t e = 0; e < (size_t)d.dim * d.dim; e++) {
    md[2 * e] = d.m[e].real();
    md[2 * e + 1] = d.m[e].imag();
  }
  if (f64) {
    ALEPH_HIP_CHECK(hipMemcpyAsync(scratch, md.data(), md.size() * sizeof(double), hipMemcpyHostToDevice, s));
    launch_kq<double>(s, state, n, k, d.qubits.data(), cm, scratch);
  } else {
    std::vector<float> mf(md.begin(), md.end());
    ALEPH_HIP_CHECK(hipMemcpyAsync(scratch, mf.data(), mf.size() * sizeof(float), hipMemcpyHostToDevice, s));
    launch_kq<float>(s, state, n, k, d.qubits.data(), cm, scratch);
  }
  ALEPH_HIP_CHECK(hipStreamSynchronize(s));
}

}  // namespace aleph
