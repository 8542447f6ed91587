// Host side of batched execution (see batch.hpp and k_batch.hip).
#include "batch.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <exception>
#include <mutex>
#include <numeric>
#include <thread>

#include "gates.hpp"
#include "jit.hpp"

namespace aleph {

namespace {

constexpr uint32_t kBatchRegBits = 4;
constexpr uint32_t kDefaultChunksPerItem = 16;
constexpr uint32_t kDefaultSpinLimit = 1u << 20;

DenseGate make_gate(uint32_t kind, std::initializer_list<uint32_t> qs, double theta, uint32_t n) {
  std::vector<uint32_t> q(qs);
  aleph_hip_gate g{};
  g.kind = kind;
  g.n_qubits = (uint32_t)q.size();
  g.qubits = q.data();
  g.params[0] = theta;
  return validate_gate(g, n);
}

ProgInstr gate_instr(DenseGate d) {
  ProgInstr pi;
  pi.type = ProgInstr::GATE;
  pi.gate = std::move(d);
  return pi;
}

template <typename V>
void ensure_upload(const std::vector<V>& v, V** dst, size_t* cap, hipStream_t s, bool* realloc) {
  const size_t need = std::max<size_t>(v.size(), 1);
  if (*dst == nullptr || *cap < need) {
    if (*dst) ALEPH_HIP_CHECK(hipFree(*dst));
    ALEPH_HIP_CHECK(hipMalloc(reinterpret_cast<void**>(dst), need * sizeof(V)));
    *cap = need;
    *realloc = true;
  }
  if (!v.empty()) ALEPH_HIP_CHECK(hipMemcpyAsync(*dst, v.data(), v.size() * sizeof(V), hipMemcpyHostToDevice, s));
}

}  // namespace

std::vector<ProgInstr> qaoa_program(uint32_t n, const uint32_t* edges, uint32_t n_edges, uint32_t p,
                                    const double* gammas, const double* betas) {
  std::vector<ProgInstr> prog;
  prog.reserve(n + (size_t)p * (3 * n_edges + n));
  for (uint32_t q = 0; q < n; q++) prog.push_back(gate_instr(make_gate(ALEPH_HIP_GATE_H, {q}, 0.0, n)));
  for (uint32_t l = 0; l < p; l++) {
    for (uint32_t e = 0; e < n_edges; e++) {
      const uint32_t u = edges[2 * e], v = edges[2 * e + 1];
      prog.push_back(gate_instr(make_gate(ALEPH_HIP_GATE_CNOT, {u, v}, 0.0, n)));
      prog.push_back(gate_instr(make_gate(ALEPH_HIP_GATE_RZ, {v}, gammas[l], n)));
      prog.push_back(gate_instr(make_gate(ALEPH_HIP_GATE_CNOT, {u, v}, 0.0, n)));
    }
    for (uint32_t q = 0; q < n; q++) prog.push_back(gate_instr(make_gate(ALEPH_HIP_GATE_RX, {q}, betas[l], n)));
  }
  return prog;
}

uint64_t strip_leading_hadamards(std::vector<ProgInstr>& prog, uint32_t n) {
  const uint64_t all = n >= 64 ? ~0ull : (1ull << n) - 1ull;
  uint64_t touched = 0, plus = 0;
  size_t w = 0, i = 0;
  for (; i < prog.size() && touched != all; i++) {
    ProgInstr& pi = prog[i];
    if (pi.type == ProgInstr::GATE) {
      const DenseGate& g = pi.gate;
      if (g.kind == ALEPH_HIP_GATE_H && g.controls.empty() && g.qubits.size() == 1 &&
          !((touched >> g.qubits[0]) & 1ull)) {
        plus |= 1ull << g.qubits[0];
        touched |= 1ull << g.qubits[0];
        continue;
      }
      for (uint32_t q : g.qubits) touched |= 1ull << q;
      for (uint32_t q : g.controls) touched |= 1ull << q;
    } else if (pi.type == ProgInstr::DIAG_PHASE) {
      for (const auto& c : pi.conds)
        for (uint64_t mk : c) touched |= mk;
    }
    if (w != i) prog[w] = std::move(pi);
    w++;
  }
  for (; i < prog.size(); i++, w++)
    if (w != i) prog[w] = std::move(prog[i]);
  prog.resize(w);
  return plus;
}

CutDesc cut_desc(uint32_t n, const uint32_t* edges, uint32_t n_edges) {
  CutDesc d{};
  d.n = n;
  for (uint32_t e = 0; e < n_edges; e++) {
    const uint32_t u = edges[2 * e], v = edges[2 * e + 1];
    if (u >= n || v >= n || u == v) {
      Failure f = make_failure(ALEPH_HIP_ERR_QUBIT_OUT_OF_RANGE, "maxcut", "edge endpoint out of range or a loop");
      f.qubit = std::max(u, v);
      f.num_qubits = n;
      throw f;
    }
    d.adj[u] |= 1u << v;
    d.adj[v] |= 1u << u;
  }
  uint64_t ll = 0;
  for (uint32_t r = 0; r < 16; r++) {
    uint32_t cut = 0;
    for (uint32_t e = 0; e < n_edges; e++) {
      const uint32_t u = edges[2 * e], v = edges[2 * e + 1];
      if (u < 4 && v < 4 && (((r >> u) ^ (r >> v)) & 1u)) cut++;
    }
    ll |= (uint64_t)cut << (4 * r);
  }
  d.ll = ll;
  return d;
}

Batch::~Batch() {
  drop_graph();
  for (void* p : {(void*)pool, (void*)d_passes, (void*)d_stages, (void*)d_ops, (void*)d_terms, (void*)d_items,
                  (void*)d_offs, (void*)d_prep, (void*)d_done, (void*)d_ctl})
    if (p) (void)hipFree(p);
}

void Batch::drop_graph() {
  if (graph) (void)hipGraphExecDestroy(graph);
  if (graph_src) (void)hipGraphDestroy(graph_src);
  graph = nullptr;
  graph_src = nullptr;
}

void Batch::plan() {
  const auto t0 = std::chrono::steady_clock::now();
  const uint32_t N = size();
  const uint32_t m_cfg = cfg.local_qubits ? cfg.local_qubits : 12u;
  const uint32_t solo_q = cfg.solo_qubits ? cfg.solo_qubits : kDefaultSoloQubits;
  if (m_cfg < kBatchRegBits + 1 || m_cfg > (uint32_t)kMaxLocalBits) {
    throw make_failure(ALEPH_HIP_ERR_INVALID_ARGUMENT, "batch", "local_qubits must be in [5, 13]");
  }
  std::vector<std::unique_ptr<Plan>> plans(N);
  plus_mask.assign(N, 0);
  std::mutex mu;
  std::exception_ptr first_error;
  std::atomic<uint32_t> next{0};
  auto worker = [&] {
    for (;;) {
      const uint32_t i = next.fetch_add(1);
      if (i >= N) return;
      try {
        PlanOptions po;
        po.planner = ALEPH_HIP_PLANNER_RULE;
        po.local_qubits = std::min(m_cfg, nq[i]);
        po.max_ops = cfg.max_ops;
        po.reg_bits = kBatchRegBits;
        po.nontemporal = cfg.nontemporal;
        const bool own = nq[i] >= solo_q;
        if (own) {
          po.local_qubits = 0;        // the single-state planner's chunking
          po.jit = true;
          po.generic_phase = true;    // instances of one size share their pass kernels
        }
        std::vector<ProgInstr> prog;
        if (qaoa.empty()) {
          prog = progs[i];
        } else {
          const QaoaSpec& q = qaoa[i];
          prog = qaoa_program(nq[i], q.edges.data(), (uint32_t)(q.edges.size() / 2), (uint32_t)q.gammas.size(),
                              q.gammas.data(), q.betas.data());
        }
        plus_mask[i] = strip_leading_hadamards(prog, nq[i]);
        std::unique_ptr<Plan> p = build_plan(prog, nq[i], precision, po);
        for (const Step& s : p->steps)
          if (!own && s.kind != Step::FUSED)
            throw make_failure(ALEPH_HIP_ERR_UNSUPPORTED_INSTRUCTION, "batch",
                               "batched circuits must plan into fused passes only (no wide dense blocks or "
                               "wide parity terms)");
        plans[i] = std::move(p);
      } catch (...) {
        std::lock_guard<std::mutex> lk(mu);
        if (!first_error) first_error = std::current_exception();
        next.store(N);
        return;
      }
    }
  };
  const uint32_t hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  const uint32_t nthreads = std::min(hw, std::max(1u, N));
  std::vector<std::thread> pool_threads;
  for (uint32_t t = 1; t < nthreads; t++) pool_threads.emplace_back(worker);
  worker();
  for (auto& t : pool_threads) t.join();
  if (first_error) std::rethrow_exception(first_error);

  passes.clear();
  stages.clear();
  ops.clear();
  terms.clear();
  offs.assign(N, 0);
  inst_passes.assign(N, 0);
  std::vector<std::vector<uint32_t>> pass_list(N);
  solo.clear();
  solo_inst.clear();
  uint64_t amp = 0;
  uint32_t m_max = kBatchRegBits;
  const double bpa = precision == ALEPH_HIP_F64 ? 16.0 : 8.0;
  sweep_bytes = 0.0;
  for (uint32_t i = 0; i < N; i++) {
    offs[i] = amp;
    amp += 1ull << nq[i];
    if (nq[i] >= solo_q) {
      inst_passes[i] = (uint32_t)plans[i]->steps.size();
      sweep_bytes += (double)inst_passes[i] * 2.0 * std::ldexp(bpa, (int)nq[i]);
      solo_inst.push_back(i);
      solo.push_back(std::move(plans[i]));
      continue;
    }
    const Plan& p = *plans[i];
    const uint32_t sb = (uint32_t)stages.size(), ob = (uint32_t)ops.size(), tb = (uint32_t)terms.size();
    for (const StageDev& sd : p.stages) {
      StageDev c = sd;
      c.op_begin += ob;
      c.op_end += ob;
      stages.push_back(c);
    }
    for (const DevOp& op : p.ops) {
      DevOp c = op;
      if (c.type == OP_PHASE) c.term_off += tb;
      ops.push_back(c);
    }
    terms.insert(terms.end(), p.terms.begin(), p.terms.end());
    for (const Step& s : p.steps) {
      PassDev pd = p.passes[s.pass];
      pd.stage_begin += sb;
      m_max = std::max(m_max, pd.m);
      pass_list[i].push_back((uint32_t)passes.size());
      passes.push_back(pd);
    }
    inst_passes[i] = (uint32_t)pass_list[i].size();
    sweep_bytes += (double)inst_passes[i] * 2.0 * std::ldexp(bpa, (int)nq[i]);
  }
  total_amps = amp;
  total_passes = (uint32_t)passes.size();
  for (const auto& sp : solo) total_passes += (uint32_t)sp->steps.size();
  if (!solo.empty()) solo_generation++;

  // Items, round-major; within a round the largest instances first.
  std::vector<uint32_t> order(N);
  std::iota(order.begin(), order.end(), 0u);
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return nq[a] > nq[b]; });
  const uint32_t cpi = cfg.chunks_per_item ? cfg.chunks_per_item : kDefaultChunksPerItem;
  uint32_t rounds = 0;
  for (uint32_t i = 0; i < N; i++)
    if (nq[i] < solo_q) rounds = std::max(rounds, inst_passes[i]);
  items.clear();
  round_begin.clear();
  std::vector<uint32_t> dep(N, 0);
  for (uint32_t r = 0; r < rounds; r++) {
    round_begin.push_back((uint32_t)items.size());
    for (uint32_t i : order) {
      if (nq[i] >= solo_q || r >= inst_passes[i]) continue;
      const uint32_t g = pass_list[i][r];
      const uint32_t chunks = 1u << (nq[i] - passes[g].m);
      for (uint32_t c = 0; c < chunks; c += cpi) {
        BatchItem it{};
        it.state_off = offs[i];
        it.pass = g;
        it.inst = i;
        it.chunk_begin = c;
        it.chunk_count = std::min(cpi, chunks - c);
        it.dep_target = dep[i] + (dep[i] != 0 && cfg.unreachable_deps ? (1u << 30) : 0u);
        items.push_back(it);
      }
      dep[i] += chunks;
    }
  }
  round_begin.push_back((uint32_t)items.size());

  launch.threads = std::max(256u, 1u << (m_max - kBatchRegBits));
  if (launch.threads > 512) throw make_failure(ALEPH_HIP_ERR_INVALID_ARGUMENT, "batch", "chunk too large");
  launch.lds_bytes = (size_t)(precision == ALEPH_HIP_F64 ? 16 : 8) << m_max;
  launch.nontemporal = cfg.nontemporal;
  plan_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

void Batch::upload(hipStream_t s, int device) {
  const size_t bpa = precision == ALEPH_HIP_F64 ? 16 : 8;
  const size_t need = (size_t)total_amps * bpa;
  if (pool == nullptr || pool_bytes < need) {
    if (pool) {
      ALEPH_HIP_CHECK(hipFree(pool));
      pool = nullptr;
    }
    size_t free_b = 0, total_b = 0;
    ALEPH_HIP_CHECK(hipMemGetInfo(&free_b, &total_b));
    if (need + (256u << 20) > free_b) {
      Failure f = make_failure(ALEPH_HIP_ERR_OUT_OF_MEMORY, "batch", "batch states do not fit in device memory");
      f.value = (double)need;
      throw f;
    }
    ALEPH_HIP_CHECK(hipMalloc(&pool, need));
    pool_bytes = need;
    drop_graph();
  }
  bool realloc = false;
  ensure_upload(passes, &d_passes, &cap_passes, s, &realloc);
  ensure_upload(stages, &d_stages, &cap_stages, s, &realloc);
  ensure_upload(ops, &d_ops, &cap_ops, s, &realloc);
  ensure_upload(terms, &d_terms, &cap_terms, s, &realloc);
  ensure_upload(items, &d_items, &cap_items, s, &realloc);
  const uint32_t N = size();
  if (!d_offs) {
    ALEPH_HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&d_offs), std::max<size_t>(N, 1) * sizeof(uint64_t)));
    ALEPH_HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&d_done), std::max<size_t>(N, 1) * sizeof(uint32_t)));
    ALEPH_HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&d_ctl), sizeof(BatchCtl)));
  }
  if (N) ALEPH_HIP_CHECK(hipMemcpyAsync(d_offs, offs.data(), N * sizeof(uint64_t), hipMemcpyHostToDevice, s));
  prep.resize(N);
  for (uint32_t i = 0; i < N; i++) {
    prep[i] = BatchPrep{};
    prep[i].state_off = offs[i];
    prep[i].plus = plus_mask[i];
    prep[i].amp = std::ldexp(1.0, -(int)(__builtin_popcountll(plus_mask[i]) / 2));
    if (__builtin_popcountll(plus_mask[i]) & 1) prep[i].amp *= std::sqrt(0.5);
    prep[i].n = nq[i];
  }
  ensure_upload(prep, &d_prep, &cap_prep, s, &realloc);
  if (realloc) drop_graph();
  for (auto& sp : solo) {
    if (!sp->uploaded) sp->upload(s);
    if (sp->jit.empty()) jit_compile_plan(*sp);
  }
  int cus = 0;
  ALEPH_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
  const int occ = precision == ALEPH_HIP_F64 ? batch_persistent_occupancy<double>(launch)
                                             : batch_persistent_occupancy<float>(launch);
  if (occ < 1) throw make_failure(ALEPH_HIP_ERR_DEVICE_FAILURE, "batch", "persistent executor cannot be resident");
  const uint32_t per_cu = cfg.groups_per_cu ? std::min<uint32_t>(cfg.groups_per_cu, (uint32_t)occ) : (uint32_t)occ;
  grid = (uint32_t)std::min<uint64_t>((uint64_t)per_cu * (uint32_t)cus, std::max<size_t>(items.size(), 1));
  if (precision == ALEPH_HIP_F64) batch_configure<double>(launch);
  else batch_configure<float>(launch);
  ALEPH_HIP_CHECK(hipStreamSynchronize(s));
}

namespace {
template <typename T>
void record_batch(Batch& b, hipStream_t s) {
  const uint32_t N = b.size();
  launch_batch_init<T>(s, b.pool, b.d_prep, N, b.d_done, b.d_ctl);
  if (b.cfg.mode == BATCH_PERSISTENT) {
    launch_batch_persistent<T>(s, b.launch, b.pool, b.d_items, (uint32_t)b.items.size(), b.grid, b.d_passes,
                               b.d_stages, b.d_ops, b.d_terms, b.d_done, b.d_ctl,
                               b.cfg.spin_limit ? b.cfg.spin_limit : kDefaultSpinLimit);
  } else {
    for (size_t r = 0; r + 1 < b.round_begin.size(); r++)
      launch_batch_round<T>(s, b.launch, b.pool, b.d_items + b.round_begin[r], b.round_begin[r + 1] - b.round_begin[r],
                            b.d_passes, b.d_stages, b.d_ops, b.d_terms);
  }
  for (size_t k = 0; k < b.solo.size(); k++)
    b.solo[k]->execute(s, static_cast<T*>(b.pool) + 2 * b.offs[b.solo_inst[k]]);
}
}  // namespace

Batch::GraphKey Batch::current_key() const {
  GraphKey k;
  k.mode = cfg.mode;
  k.grid = grid;
  k.threads = launch.threads;
  k.n_items = (uint32_t)items.size();
  k.solo_generation = solo_generation;
  k.lds_bytes = launch.lds_bytes;
  k.nontemporal = launch.nontemporal;
  k.round_begin = round_begin;
  k.ptrs = {pool, d_passes, d_stages, d_ops, d_terms, d_items, d_offs, d_prep, d_done, d_ctl};
  return k;
}

void Batch::run(hipStream_t s) {
  GraphKey key = current_key();
  if (graph && !(key == graph_key)) drop_graph();
  if (!graph) {
    ALEPH_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    try {
      if (precision == ALEPH_HIP_F64) record_batch<double>(*this, s);
      else record_batch<float>(*this, s);
    } catch (...) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(s, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    ALEPH_HIP_CHECK(hipStreamEndCapture(s, &graph_src));
    ALEPH_HIP_CHECK(hipGraphInstantiate(&graph, graph_src, nullptr, nullptr, 0));
    graph_key = std::move(key);
  }
  ALEPH_HIP_CHECK(hipGraphLaunch(graph, s));
}

bool Batch::failed(hipStream_t s) {
  if (cfg.mode != BATCH_PERSISTENT || !d_ctl) return false;
  BatchCtl h{};
  ALEPH_HIP_CHECK(hipMemcpyAsync(&h, d_ctl, sizeof(h), hipMemcpyDeviceToHost, s));
  ALEPH_HIP_CHECK(hipStreamSynchronize(s));
  return h.error != 0;
}

}  // namespace aleph
