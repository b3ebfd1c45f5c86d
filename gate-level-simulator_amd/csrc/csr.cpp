#include "csr.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <exception>
#include <map>
#include <thread>
#include <unordered_map>
#include <unordered_set>

#include "../../include/gls_status.h"

namespace gls {

static inline bool is01(uint8_t v) { return v == V0 || v == V1; }
static inline bool isXZ(uint8_t v) { return v == VX || v == VZ; }
static inline uint8_t inv(uint8_t v) { return v == V0 ? V1 : (v == V1 ? V0 : VX); }

// and/or/xor over the first two inputs, then folded over the rest, with the
// reference's case analysis (src/simulate.cpp:384-823).
static uint8_t and_n(const uint8_t* in, int n) {
  uint8_t t;
  if (is01(in[0]) && is01(in[1])) t = (in[0] == V1 && in[1] == V1) ? V1 : V0;
  else if (isXZ(in[0])) t = in[1] == V0 ? V0 : VX;
  else t = in[0] == V0 ? V0 : VX;
  for (int i = 2; i < n; ++i) {
    if (t == V0) t = V0;
    else if (t == V1) t = in[i] == VZ ? VX : in[i];
    else t = in[i] == V0 ? V0 : VX;
  }
  return t;
}
static uint8_t or_n(const uint8_t* in, int n) {
  uint8_t t;
  if (is01(in[0]) && is01(in[1])) t = (in[0] == V1 || in[1] == V1) ? V1 : V0;
  else if (isXZ(in[0])) t = in[1] == V1 ? V1 : VX;
  else t = in[0] == V1 ? V1 : VX;
  for (int i = 2; i < n; ++i) {
    if (t == V1) t = V1;
    else if (t == V0) t = in[i] == VZ ? VX : in[i];
    else t = in[i] == V1 ? V1 : VX;
  }
  return t;
}
static uint8_t xor_n(const uint8_t* in, int n) {
  uint8_t t;
  if (is01(in[0]) && is01(in[1])) t = (in[0] != in[1]) ? V1 : V0;
  else if (in[0] == VX || in[1] == VX) t = VX;
  else t = (in[0] == VZ && in[1] == VZ) ? V0 : VX;  // Z^Z = 0 (src/simulate.cpp:545-551)
  for (int i = 2; i < n; ++i) {
    if (t == V0) t = in[i] == VZ ? VX : in[i];
    else if (t == V1) t = is01(in[i]) ? inv(in[i]) : VX;
    else t = VX;
  }
  return t;
}

uint8_t eval_prim(int op, const uint8_t* in, int n) {
  switch (op) {
    case P_AND: return and_n(in, n);
    case P_OR: return or_n(in, n);
    case P_XOR: return xor_n(in, n);
    case P_XNOR: return inv(xor_n(in, n));
    case P_NOR: return inv(or_n(in, n));
    case P_NAND: return inv(and_n(in, n));
    case P_BUF: return in[0] == VZ ? VX : in[0];
    case P_NOT: return inv(in[0]);
    case P_XBUF: return in[1] == V1 ? (in[0] == VX ? V1 : in[0]) : VX;
    case P_MUX2:
      if (in[2] == V0) return in[0];
      if (in[2] == V1) return in[1];
      if (in[2] == VX) return in[0] == in[1] ? in[0] : VX;
      return VX;
  }
  return VX;
}

namespace {

// Per-instance name tables used to evaluate a cell's function chain with
// the reference's name lookups (src/simulate.cpp:1203-1363): primitive pins
// bound to an instance input are renamed to the connected net; every other
// pin keeps its module-local name and is looked up, in order, among the
// instance's input nets, earlier primitive outputs, supply1, supply0,
// else X.
struct ChainEval {
  struct PrimRef { int op; std::string out; std::vector<std::string> ins; };
  std::vector<PrimRef> prims;
  std::vector<std::string> in_names;  // distinct input net names (index j)
  std::vector<std::string> s0, s1;

  bool run(const std::string& target, const uint8_t* vals, uint8_t& out) const {
    std::unordered_map<std::string, uint8_t> inter;
    uint8_t buf[16];
    for (auto& p : prims) {
      int n = 0;
      for (auto& nm : p.ins) {
        uint8_t v = VX;
        bool found = false;
        for (size_t j = 0; j < in_names.size(); ++j)
          if (in_names[j] == nm) { v = vals[j]; found = true; break; }
        if (!found) {
          auto it = inter.find(nm);
          if (it != inter.end()) { v = it->second; found = true; }
        }
        if (!found && std::find(s1.begin(), s1.end(), nm) != s1.end()) { v = V1; found = true; }
        if (!found && std::find(s0.begin(), s0.end(), nm) != s0.end()) { v = V0; found = true; }
        if (n >= 16) return false;
        buf[n++] = v;
      }
      int need = (p.op == P_BUF || p.op == P_NOT) ? 1 : (p.op == P_XBUF ? 2 : (p.op == P_MUX2 ? 3 : 2));
      if (n < need) return false;
      uint8_t r = eval_prim(p.op, buf, n);
      if (p.out == target) { out = r; return true; }
      inter[p.out] = r;
    }
    return false;
  }
};

}  // namespace

static int delay_kind(int edge) { return edge == 2 ? DK_BOTH : (edge == 0 ? DK_POS : DK_NEG); }

namespace {

// Gates of a contiguous run of instances, compiled by one thread exactly as
// a serial pass would compile them, with numbers local to the run: truth
// tables index `luts` (first appearance in the run), and the pseudo nets of
// unconnected outputs are numbered N + their index in `pseudo`.
struct CompiledRun {
  std::vector<CompiledGate> gates;
  std::vector<std::vector<uint8_t>> luts;
  std::vector<int> lut_k;
  std::vector<std::string> pseudo;
  std::exception_ptr err;  // the run stopped at its first error
};

void compile_run(const Netlist& nl, const Library& lib, const std::vector<uint8_t>& role,
                 size_t i0, size_t i1, CompiledRun& R) {
  const int N = (int)nl.net_names.size();
  // resolve a net read by an instance input to the net whose waveform is used
  auto resolve = [&](int n) -> int {
    if (role[n] == NR_ALIAS) {
      int t = nl.alias_of[n];  // one level, as the reference does
      if (role[t] == NR_STIM || role[t] == NR_GATE || role[t] == NR_CONST) return t;
      return -1;
    }
    if (role[n] == NR_UNDRIVEN) return -1;
    return n;
  };
  std::map<std::string, int> lut_cache;  // signature -> local lut id
  R.gates.reserve(i1 - i0);
  for (size_t ii = i0; ii < i1; ++ii) {
    const Instance& inst = nl.insts[ii];
    const CellType& ct = lib.types[inst.type];
    std::vector<int> dnets = inst.in_nets();
    if ((int)dnets.size() > kMaxK)
      throw GlsError(GLS_E_UNSUPPORTED, "instance " + inst.name + " has more than 6 distinct input nets");
    std::vector<int> rnets(dnets.size());
    for (size_t j = 0; j < dnets.size(); ++j) {
      rnets[j] = resolve(dnets[j]);
      if (rnets[j] < 0)
        throw GlsError(GLS_E_CYCLE, "net " + nl.net_names[dnets[j]] + " read by " + inst.name + " has no driver");
    }
    // chain with pin -> net renaming
    ChainEval ce;
    for (int d : dnets) ce.in_names.push_back(nl.net_names[d]);
    ce.s0 = ct.supply0;
    ce.s1 = ct.supply1;
    std::string sig = std::to_string(inst.type) + ":";
    bool collide = false;
    for (auto& p : ct.prims) {
      ChainEval::PrimRef pr;
      pr.op = p.op;
      pr.out = p.out;
      for (auto& op : inst.out_pins)
        if (op.first == p.out) pr.out = nl.net_names[op.second];
      for (auto& in : p.ins) {
        std::string nm = in;
        for (auto& ip : inst.in_pins)
          if (ip.first == in) { nm = nl.net_names[ip.second]; break; }
        if (nm == in) {  // module-local name: does it collide with an input net?
          for (auto& d : ce.in_names) collide |= (d == nm);
        }
        pr.ins.push_back(nm);
        // signature: position of the net among the distinct inputs, or the local name
        int pos = -1;
        for (size_t j = 0; j < ce.in_names.size(); ++j)
          if (ce.in_names[j] == nm) pos = (int)j;
        sig += pos >= 0 ? ("#" + std::to_string(pos)) : ("@" + nm);
        sig += ",";
      }
      sig += ";";
      ce.prims.push_back(pr);
    }
    struct OutSpec { std::string target; int net; std::string pin; };
    std::vector<OutSpec> outs;
    for (auto& op : inst.out_pins) outs.push_back({nl.net_names[op.second], op.second, op.first});
    for (size_t k = 0; k < inst.unknown_out.size(); ++k) {
      const int id = N + (int)R.pseudo.size();
      R.pseudo.push_back(inst.name + "[" + inst.unknown_out[k] + "]");
      outs.push_back({inst.unknown_out[k], id, inst.unknown_out[k]});
    }
    for (size_t oi = 0; oi < outs.size(); ++oi) {
      const OutSpec& os = outs[oi];
      if (os.net < N && role[os.net] != NR_GATE) continue;  // a VCD signal overrides this output
      CompiledGate g;
      g.inst = (int)ii;
      g.out_net = os.net;
      g.k = (int)dnets.size();
      for (int j = 0; j < g.k; ++j) g.in_net[j] = rnets[j];
      // truth table
      std::string key = sig + "|" + os.pin + "|" + std::to_string(oi < inst.out_pins.size() ? 0 : 1);
      auto it = collide ? lut_cache.end() : lut_cache.find(key);
      if (it != lut_cache.end()) {
        g.lut = it->second;
      } else {
        int ent = 1 << (2 * g.k);
        std::vector<uint8_t> t(ent);
        uint8_t vals[kMaxK];
        for (int e = 0; e < ent; ++e) {
          for (int j = 0; j < g.k; ++j) vals[j] = (uint8_t)((e >> (2 * j)) & 3);
          uint8_t r;
          if (!ce.run(os.target, vals, r))
            throw GlsError(GLS_E_UNSUPPORTED, "no primitive drives output " + os.pin + " of " + inst.name);
          t[e] = r;
        }
        int id = -1;
        for (size_t q = 0; q < R.luts.size(); ++q)
          if (R.luts[q] == t && R.lut_k[q] == g.k) { id = (int)q; break; }
        if (id < 0) {
          id = (int)R.luts.size();
          R.luts.push_back(t);
          R.lut_k.push_back(g.k);
        }
        if (!collide) lut_cache[key] = id;
        g.lut = id;
      }
      // delays related to this output, grouped per distinct input net
      int want_out = (oi < inst.out_pins.size()) ? inst.out_pins[oi].second : -2 - (int)(oi - inst.out_pins.size());
      for (auto& d : inst.delays) {
        if (d.out_net != want_out) continue;
        for (int j = 0; j < g.k; ++j)
          if (d.in_net == dnets[j]) g.recs[j].push_back({delay_kind(d.edge), d.rise, d.fall});
      }
      for (int j = 0; j < g.k; ++j) {
        int seen[3] = {0, 0, 0};
        for (auto& r : g.recs[j]) {
          if (++seen[r.kind] > 1)
            throw GlsError(GLS_E_UNSUPPORTED, "instance " + inst.name + " has duplicate IOPATHs of one kind on one input net");
          if (!(r.rise >= 0) || !(r.fall >= 0))
            throw GlsError(GLS_E_UNSUPPORTED, "negative delay on " + inst.name);
        }
      }
      R.gates.push_back(std::move(g));
    }
  }
}

}  // namespace

Program compile_program(const Netlist& nl, const Library& lib,
                        const std::vector<std::string>& stim_names, int threads) {
  Program pg;
  const int N = (int)nl.net_names.size();
  pg.net_names = nl.net_names;
  pg.n_netlist_nets = N;
  pg.role.assign(N, NR_UNDRIVEN);
  pg.const_val.assign(N, -1);
  pg.in_saif.assign(N, 0);

  std::unordered_set<std::string> stim(stim_names.begin(), stim_names.end());
  // gate outputs
  std::vector<int> driver_gate(N, -1);  // temporary: instance index
  for (size_t i = 0; i < nl.insts.size(); ++i)
    for (auto& op : nl.insts[i].out_pins) driver_gate[op.second] = (int)i;
  for (int n = 0; n < N; ++n) {
    if (stim.count(nl.net_names[n])) pg.role[n] = NR_STIM;
    else if (driver_gate[n] >= 0) pg.role[n] = NR_GATE;
    else if (nl.const_val[n] >= 0) { pg.role[n] = NR_CONST; pg.const_val[n] = nl.const_val[n]; }
    else if (nl.alias_of[n] >= 0) pg.role[n] = NR_ALIAS;
  }

  // Compile gates (one per connected output and per unknown_out) in runs of
  // instances on parallel threads, then number truth tables and pseudo nets
  // in instance order: the first appearance of a table in the netlist fixes
  // its number, exactly as one serial pass numbers it.
  const size_t NI = nl.insts.size();
  threads = host_threads(threads);
  const size_t nrun = std::max<size_t>(1, std::min<size_t>((size_t)threads, NI / 16384));
  std::vector<CompiledRun> runs(nrun);
  auto work = [&](size_t r) {
    try {
      compile_run(nl, lib, pg.role, NI * r / nrun, NI * (r + 1) / nrun, runs[r]);
    } catch (...) {
      runs[r].err = std::current_exception();
    }
  };
  if (nrun == 1) {
    work(0);
  } else {
    std::vector<std::thread> th;
    for (size_t r = 0; r < nrun; ++r) th.emplace_back(work, r);
    for (auto& t : th) t.join();
  }
  // the first error in instance order is the one a serial pass reports
  for (auto& R : runs)
    if (R.err) std::rethrow_exception(R.err);
  std::vector<CompiledGate> gates;
  size_t total = 0;
  for (auto& R : runs) total += R.gates.size();
  gates.reserve(total);
  for (auto& R : runs) {
    std::vector<int> lut_id(R.luts.size());
    for (size_t q = 0; q < R.luts.size(); ++q) {
      int id = -1;
      for (size_t p = 0; p < pg.luts.size(); ++p)
        if (pg.luts[p] == R.luts[q] && pg.lut_k[p] == R.lut_k[q]) { id = (int)p; break; }
      if (id < 0) {
        id = (int)pg.luts.size();
        pg.luts.push_back(R.luts[q]);
        pg.lut_k.push_back(R.lut_k[q]);
      }
      lut_id[q] = id;
    }
    const int pseudo0 = (int)pg.net_names.size();
    for (auto& nm : R.pseudo) {
      pg.net_names.push_back(std::move(nm));
      pg.role.push_back(NR_GATE);
      pg.const_val.push_back(-1);
      pg.in_saif.push_back(1);
    }
    for (auto& g : R.gates) {
      g.lut = lut_id[g.lut];
      if (g.out_net >= N) g.out_net = pseudo0 + (g.out_net - N);
      gates.push_back(std::move(g));
    }
    R = CompiledRun();
  }

  // SAIF membership: gate outputs, VCD signals, visible constants.
  for (int n = 0; n < N; ++n) {
    if (pg.role[n] == NR_GATE || pg.role[n] == NR_STIM) pg.in_saif[n] = 1;
    if (pg.role[n] == NR_CONST && !nl.const_hidden[n]) pg.in_saif[n] = 1;
  }

  // Levels: longest path from stimulus / constants (Kahn), readers in CSR.
  const int G = (int)gates.size();
  std::vector<int> gate_of_net(pg.net_names.size(), -1);
  for (int g = 0; g < G; ++g) gate_of_net[gates[g].out_net] = g;
  std::vector<int> indeg(G, 0);
  std::vector<uint32_t> rd_off(G + 1, 0);
  auto src_of = [&](int g, int j) {
    const int n = gates[g].in_net[j];
    const int s = gate_of_net[n];
    return (s >= 0 && pg.role[n] == NR_GATE) ? s : -1;
  };
  for (int g = 0; g < G; ++g)
    for (int j = 0; j < gates[g].k; ++j) {
      const int s = src_of(g, j);
      if (s >= 0) {
        rd_off[s + 1]++;
        indeg[g]++;
      }
    }
  for (int g = 0; g < G; ++g) rd_off[g + 1] += rd_off[g];
  std::vector<int> readers(rd_off[G]);
  {
    std::vector<uint32_t> fill(rd_off.begin(), rd_off.end() - 1);
    for (int g = 0; g < G; ++g)
      for (int j = 0; j < gates[g].k; ++j) {
        const int s = src_of(g, j);
        if (s >= 0) readers[fill[s]++] = g;
      }
  }
  std::vector<int> frontier;
  for (int g = 0; g < G; ++g)
    if (indeg[g] == 0) frontier.push_back(g);
  std::vector<std::vector<int>> by_level;
  while (!frontier.empty()) {
    std::sort(frontier.begin(), frontier.end());
    std::vector<int> next;
    for (int g : frontier)
      for (uint32_t q = rd_off[g]; q < rd_off[g + 1]; ++q)
        if (--indeg[readers[q]] == 0) next.push_back(readers[q]);
    by_level.push_back(std::move(frontier));
    frontier.swap(next);
  }
  size_t placed = 0;
  for (auto& l : by_level) placed += l.size();
  if ((int)placed != G) throw GlsError(GLS_E_CYCLE, "combinational loop in the netlist");
  pg.gates.reserve(G);
  pg.level_start.push_back(0);
  for (auto& l : by_level) {
    for (int g : l) {
      pg.n_pins += (uint64_t)gates[g].k;
      pg.gates.push_back(std::move(gates[g]));
    }
    pg.level_start.push_back((uint32_t)pg.gates.size());
  }
  return pg;
}

uint64_t program_digest(const Program& pg) {
  uint64_t h = 1469598103934665603ull;
  auto bytes = [&](const void* p, size_t n) {
    const unsigned char* b = (const unsigned char*)p;
    for (size_t i = 0; i < n; ++i) h = (h ^ b[i]) * 1099511628211ull;
  };
  auto u64 = [&](uint64_t x) { bytes(&x, 8); };
  auto f64 = [&](double x) {
    uint64_t b;
    std::memcpy(&b, &x, 8);
    u64(b);
  };
  u64(pg.net_names.size());
  for (size_t n = 0; n < pg.net_names.size(); ++n) {
    u64(pg.net_names[n].size());
    bytes(pg.net_names[n].data(), pg.net_names[n].size());
    u64(pg.role[n]);
    u64((uint64_t)(int64_t)pg.const_val[n]);
    u64(pg.in_saif[n]);
  }
  u64(pg.gates.size());
  for (const CompiledGate& g : pg.gates) {
    u64((uint64_t)(int64_t)g.inst);
    u64((uint64_t)(int64_t)g.out_net);
    u64((uint64_t)(int64_t)g.k);
    for (int j = 0; j < kMaxK; ++j) u64((uint64_t)(int64_t)g.in_net[j]);
    u64((uint64_t)(int64_t)g.lut);
    for (int j = 0; j < kMaxK; ++j) {
      u64(g.recs[j].size());
      for (auto& r : g.recs[j]) {
        u64((uint64_t)(int64_t)r.kind);
        f64(r.rise);
        f64(r.fall);
      }
    }
  }
  for (uint32_t l : pg.level_start) u64(l);
  for (size_t q = 0; q < pg.luts.size(); ++q) {
    u64((uint64_t)(int64_t)pg.lut_k[q]);
    u64(pg.luts[q].size());
    bytes(pg.luts[q].data(), pg.luts[q].size());
  }
  u64(pg.n_pins);
  u64((uint64_t)(int64_t)pg.n_netlist_nets);
  return h;
}

GateDelays gate_delays(const CompiledGate& g, int timescale_ps) {
  double factor = 1.0;
  if (timescale_ps == 1) factor = 1.0;
  else if (timescale_ps == 1000) factor = 1e-3;
  else if (timescale_ps == 1000000) factor = 1e-6;
  else throw GlsError(GLS_E_UNSUPPORTED, "VCD timescale must be 1ps, 1ns or 1us");
  GateDelays d;
  uint32_t dmax = 1;
  for (int j = 0; j < kMaxK; ++j) {
    d.nrec[j] = 0;
    for (int k = 0; k < 3; ++k) d.v[j][k][0] = d.v[j][k][1] = kNoDelay;
  }
  for (int j = 0; j < g.k; ++j) {
    d.nrec[j] = (uint8_t)g.recs[j].size();
    for (auto& r : g.recs[j]) {
      double a = r.rise * factor, b = r.fall * factor;
      if (a >= 65535.0 || b >= 65535.0)
        throw GlsError(GLS_E_UNSUPPORTED, "IOPATH delay of 65535 time units or more");
      uint16_t ia = (uint16_t)(unsigned long long)a, ib = (uint16_t)(unsigned long long)b;
      d.v[j][r.kind][0] = ia;
      d.v[j][r.kind][1] = ib;
      dmax = std::max<uint32_t>(dmax, std::max(ia, ib));
    }
  }
  d.dmax = (uint16_t)dmax;
  return d;
}

}  // namespace gls
