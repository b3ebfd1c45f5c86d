#include "csr.h"

#include <algorithm>
#include <cmath>
#include <map>
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

Program compile_program(const Netlist& nl, const Library& lib,
                        const std::vector<std::string>& stim_names) {
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
  // resolve a net read by an instance input to the net whose waveform is used
  auto resolve = [&](int n) -> int {
    if (pg.role[n] == NR_ALIAS) {
      int t = nl.alias_of[n];  // one level, as the reference does
      if (pg.role[t] == NR_STIM || pg.role[t] == NR_GATE || pg.role[t] == NR_CONST) return t;
      return -1;
    }
    if (pg.role[n] == NR_UNDRIVEN) return -1;
    return n;
  };

  // Compile gates (one per connected output and per unknown_out).
  struct Pending { CompiledGate g; };
  std::vector<CompiledGate> gates;
  gates.reserve(nl.insts.size());
  std::map<std::string, int> lut_cache;  // signature -> lut id
  for (size_t ii = 0; ii < nl.insts.size(); ++ii) {
    const Instance& inst = nl.insts[ii];
    const CellType& ct = lib.types[inst.type];
    std::vector<int> dnets = inst.in_nets();
    if ((int)dnets.size() > kMaxK)
      throw GlsError(GLS_E_UNSUPPORTED, "instance " + inst.name + " has more than 4 distinct input nets");
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
      std::string pseudo = inst.name + "[" + inst.unknown_out[k] + "]";
      int id = (int)pg.net_names.size();
      pg.net_names.push_back(pseudo);
      pg.role.push_back(NR_GATE);
      pg.const_val.push_back(-1);
      pg.in_saif.push_back(1);
      outs.push_back({inst.unknown_out[k], id, inst.unknown_out[k]});
      (void)k;
    }
    for (size_t oi = 0; oi < outs.size(); ++oi) {
      const OutSpec& os = outs[oi];
      if (os.net < N && pg.role[os.net] != NR_GATE) continue;  // a VCD signal overrides this output
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
        for (size_t q = 0; q < pg.luts.size(); ++q)
          if (pg.luts[q] == t && pg.lut_k[q] == g.k) { id = (int)q; break; }
        if (id < 0) {
          id = (int)pg.luts.size();
          pg.luts.push_back(t);
          pg.lut_k.push_back(g.k);
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
      gates.push_back(std::move(g));
    }
  }

  // SAIF membership: gate outputs, VCD signals, visible constants.
  for (int n = 0; n < N; ++n) {
    if (pg.role[n] == NR_GATE || pg.role[n] == NR_STIM) pg.in_saif[n] = 1;
    if (pg.role[n] == NR_CONST && !nl.const_hidden[n]) pg.in_saif[n] = 1;
  }

  // Levels: longest path from stimulus / constants (Kahn).
  const int G = (int)gates.size();
  std::vector<int> gate_of_net(pg.net_names.size(), -1);
  for (int g = 0; g < G; ++g) gate_of_net[gates[g].out_net] = g;
  std::vector<int> indeg(G, 0);
  std::vector<std::vector<int>> readers(G);
  for (int g = 0; g < G; ++g) {
    for (int j = 0; j < gates[g].k; ++j) {
      int src = gate_of_net[gates[g].in_net[j]];
      if (src >= 0 && pg.role[gates[g].in_net[j]] == NR_GATE) {
        readers[src].push_back(g);
        indeg[g]++;
      }
    }
  }
  std::vector<int> level(G, 0), order;
  order.reserve(G);
  std::vector<int> frontier;
  for (int g = 0; g < G; ++g)
    if (indeg[g] == 0) frontier.push_back(g);
  int lv = 0;
  std::vector<std::vector<int>> by_level;
  while (!frontier.empty()) {
    std::sort(frontier.begin(), frontier.end());
    by_level.push_back(frontier);
    std::vector<int> next;
    for (int g : frontier) {
      level[g] = lv;
      for (int r : readers[g])
        if (--indeg[r] == 0) next.push_back(r);
    }
    frontier.swap(next);
    ++lv;
  }
  size_t placed = 0;
  for (auto& l : by_level) placed += l.size();
  if ((int)placed != G) throw GlsError(GLS_E_CYCLE, "combinational loop in the netlist");
  pg.level_start.push_back(0);
  for (auto& l : by_level) {
    for (int g : l) {
      pg.gates.push_back(gates[g]);
      for (int j = 0; j < gates[g].k; ++j) pg.n_pins++;
    }
    pg.level_start.push_back((uint32_t)pg.gates.size());
  }
  return pg;
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
