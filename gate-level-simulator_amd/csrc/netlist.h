// Host-side netlist model shared by the parsers, the synthetic generator,
// the levelizer / CSR packer and the SAIF writer.
//
// The 4-state value codes are the reference's VCDBit layout
// (src/VCDTypes.h:49-54): 0 -> 0, X -> 1, Z -> 2, 1 -> 3.
#pragma once
#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <map>
#include <stdexcept>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

namespace gls {

// Host threads of the parallel readers and the netlist compiler: `requested`
// when positive, else the GLS_HOST_THREADS environment variable when set to
// a positive count, else one per hardware thread up to 16.
inline int host_threads(int requested) {
  if (requested > 0) return requested;
  if (const char* e = std::getenv("GLS_HOST_THREADS")) {
    const int v = std::atoi(e);
    if (v > 0) return v;
  }
  return (int)std::min<unsigned>(16, std::max(1u, std::thread::hardware_concurrency()));
}

// Texts below this size are read on one thread: 8 MiB, or the
// GLS_PAR_MIN_BYTES environment variable when set to a positive count (the
// sanitizer checks lower it to run the parallel readers on small texts).
inline size_t parallel_min_bytes() {
  if (const char* e = std::getenv("GLS_PAR_MIN_BYTES")) {
    const long long v = std::atoll(e);
    if (v > 0) return (size_t)v;
  }
  return (size_t)8 << 20;
}

enum : uint8_t { V0 = 0, VX = 1, VZ = 2, V1 = 3 };

// Primitive ids follow the reference's function_map
// (src/PreProcess.cpp:1194-1197).
enum PrimOp : int {
  P_AND = 0, P_OR = 1, P_XOR = 2, P_XNOR = 3, P_NOR = 4,
  P_NAND = 5, P_BUF = 6, P_NOT = 7, P_XBUF = 8, P_MUX2 = 9
};

struct Prim {
  int op = 0;
  std::string out;               // module-local name
  std::vector<std::string> ins;  // module-local names (ports, wires, supplies)
};

struct CellType {
  std::string name;
  std::vector<std::string> ports;   // input/output/inout in declaration order
  std::vector<std::string> inputs;  // input ports in declaration order
  std::vector<std::string> outputs; // output ports in declaration order
  std::vector<Prim> prims;          // in vlib order
  std::vector<std::string> supply0, supply1;
};

struct Library {
  std::vector<CellType> types;
  std::unordered_map<std::string, int> by_name;
  std::vector<std::string> ignored;  // sequential cells the reference skips
  int find(const std::string& n) const {
    auto it = by_name.find(n);
    return it == by_name.end() ? -1 : it->second;
  }
  void add(CellType t) {
    by_name[t.name] = (int)types.size();
    types.push_back(std::move(t));
  }
};

// One SDF IOPATH record, resolved to net ids (src/VCDTypes.h:166-186).
// edge: 0 posedge, 1 negedge, 2 both.  Values are already scaled to
// picoseconds by the SDF timescale, exactly as the reference stores them.
struct IoPath {
  int edge = 2;
  int in_net = -1;
  int out_net = -1;   // net id of the output, or -2 - k for unknown_out[k]
  double rise = 1.0, fall = 1.0;
};

struct Instance {
  std::string name;
  int type = -1;
  // Pin connections in the order they appear in the gv instance statement.
  std::vector<std::pair<std::string, int>> in_pins;
  std::vector<std::pair<std::string, int>> out_pins;
  std::vector<std::string> unknown_out;  // module outputs left unconnected
  std::vector<IoPath> delays;            // SDF order
  // Distinct input nets in first-appearance order (the reference's in_net
  // may repeat a net; its simulator keys inputs by net name).
  std::vector<int> in_nets() const {
    std::vector<int> r;
    for (auto& p : in_pins) {
      bool seen = false;
      for (int x : r) seen |= (x == p.second);
      if (!seen) r.push_back(p.second);
    }
    return r;
  }
};

struct Netlist {
  std::string top = "top";
  std::vector<std::string> net_names;
  std::unordered_map<std::string, int> net_ids;
  std::vector<uint8_t> net_is_input;   // declared `input` in the gv
  std::vector<int> const_val;          // -1, or V0/V1 for 1'b0/1'b1 nets
  std::vector<uint8_t> const_hidden;   // gv_net-style constants (not in SAIF)
  std::vector<int> alias_of;           // -1, or the net an `assign` aliases
  std::vector<Instance> insts;

  int net(const std::string& n) const {
    auto it = net_ids.find(n);
    return it == net_ids.end() ? -1 : it->second;
  }
  int add_net(const std::string& n) {
    auto it = net_ids.find(n);
    if (it != net_ids.end()) return it->second;
    int id = (int)net_names.size();
    net_ids.emplace(n, id);
    net_names.push_back(n);
    net_is_input.push_back(0);
    const_val.push_back(-1);
    const_hidden.push_back(0);
    alias_of.push_back(-1);
    return id;
  }
};

// A net waveform: strictly time-increasing (time, value) entries.
struct Wave {
  std::vector<uint32_t> t;
  std::vector<uint8_t> v;
  size_t size() const { return t.size(); }
  void push(uint32_t tt, uint8_t vv) { t.push_back(tt); v.push_back(vv); }
};

// Stimulus: waveforms for the VCD signals, keyed by signal reference name.
struct Stimulus {
  int timescale_ps = 1;            // VCD time unit in ps (1 for `1ps`)
  std::string scope = "top";
  std::vector<std::string> names;  // signal names in declaration order
  std::vector<Wave> waves;         // parallel to names
};

struct GlsError : std::runtime_error {
  int code;
  GlsError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

}  // namespace gls
