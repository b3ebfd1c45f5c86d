// ORACLE — test infrastructure only.  Nothing in the product links, loads or
// calls this file; tests, __graft_entry__.smoke() and bench.py's CPU baseline
// use it as the checker.
//
// cpu_ref: a serial CPU restatement of the reference's serial simulator
// (src/simulate.cpp) with its own minimal readers for the vlib / gv / SDF /
// VCD subset the reference accepts.  It keeps the reference's per-event
// structure: per output, the input waveforms are merged into a deque of
// (time, {net -> value}) records by binary-search insertion (Process), every
// input is sampled by binary search (GetValueAt), the function chain is run
// primitive by primitive with calc(), the output edge is classified
// (CheckDelay), the SDF delay is picked from the related IOPATHs (AddDelay,
// double arithmetic) and the output queue is updated with preemption
// (Calculate).  SAIF durations follow DumpSaif's loop.
//
// Deliberate, documented choices where the reference is ill-defined:
//   * calc() is read with its intended encoding (X = 2, Z = 3 in its switch
//     comments, src/simulate.cpp:439-442) by translating the VCDBit codes
//     (src/VCDTypes.h:49-54) in and out.
//   * asserts are taken as compiled out (NDEBUG branch of AddDelay).
//   * several VCD changes of one signal at one timestamp: the last one wins.
//   * SAIF window comes from the command line (src/simulate.cu:1811-1812),
//     nets are written in lexicographic order.
//
// Parity pins for this file live in tests/test_oracle.py (truth tables vs
// the Verilog primitive definitions and the udp_mux2 table of the vlib,
// hand-computed known-answer waveforms per rule, brute force of the
// preemption queue against its closed form, SAIF closed forms).
#include <algorithm>
#include <array>
#include <cassert>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <fstream>
#include <limits>
#include <map>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

namespace oracle {

enum VCDBit { VCD_0 = 0, VCD_X = 1, VCD_Z = 2, VCD_1 = 3 };  // src/VCDTypes.h:49-54
enum TimeUnit { TIME_S, TIME_MS, TIME_US, TIME_NS, TIME_PS };  // src/VCDTypes.h:83-89

struct TimedValue { uint64_t t; VCDBit value; };
typedef std::deque<TimedValue> TimedValues;
struct TimedValueSim { uint64_t time; std::map<int, VCDBit> values; };

struct Function {  // src/Function.h:13-66 (pins already renamed to nets)
  std::string name;
  std::vector<std::string> input_pins;
  std::string output;
};
struct ModuleType {
  std::string name;
  std::vector<std::string> in_out_ports;
  std::map<std::string, std::string> port_type;
  std::vector<std::string> supply0, supply1;
  std::vector<Function> function;
};
struct Delay { int edge; int in_bit; std::string out_bit; double rise_val, fall_val; };
struct Instance {
  std::string name;
  int type;
  std::map<std::string, std::string> in_port_net, out_port_net;
  std::vector<std::string> in_net, out_net, unknown_out;
  std::vector<Function> functions;
  std::vector<Delay> delays;
};

static std::vector<std::string> Split(const std::string& s, const char* d) {
  std::vector<std::string> r;
  std::string cur;
  for (char c : s) {
    if (std::strchr(d, c) && c) {
      if (!cur.empty()) r.push_back(cur);
      cur.clear();
    } else {
      cur += c;
    }
  }
  if (!cur.empty()) r.push_back(cur);
  return r;
}
static std::string lower(std::string s) { for (auto& c : s) c = (char)tolower(c); return s; }
static std::string upper(std::string s) { for (auto& c : s) c = (char)toupper(c); return s; }
static std::string trim(const std::string& s) {
  size_t a = s.find_first_not_of(' ');
  if (a == std::string::npos) return "";
  return s.substr(a, s.find_last_not_of(' ') - a + 1);
}
static std::vector<std::string> read_lines(const std::string& path) {
  std::ifstream f(path);
  if (!f) { std::fprintf(stderr, "oracle: cannot open %s\n", path.c_str()); std::exit(3); }
  std::vector<std::string> r;
  std::string l;
  while (std::getline(f, l)) r.push_back(l);
  return r;
}

struct Design {
  std::vector<ModuleType> types;
  std::map<std::string, int> type_by_name;
  std::vector<Instance> insts;
  std::map<std::string, int> net_id;     // every declared / used net
  std::vector<std::string> net_name;
  std::map<std::string, TimedValues> pinbitValues;  // constants
  std::map<std::string, std::string> assign_pairs;
  int net(const std::string& n) {
    auto it = net_id.find(n);
    if (it != net_id.end()) return it->second;
    int id = (int)net_name.size();
    net_id[n] = id;
    net_name.push_back(n);
    return id;
  }
};

// parse_vlib, src/PreProcess.cpp:42-648 (subset)
static void parse_vlib(const std::string& path, Design& D) {
  auto L = read_lines(path);
  int zc = 0;
  for (size_t i = 0; i < L.size(); ++i) {
    auto t = Split(L[i], " \t\r\n,;()");
    if (t.empty() || lower(t[0]) != "module") continue;
    ModuleType m;
    m.name = t[1];
    bool ignored = false;
    for (++i; i < L.size(); ++i) {
      auto u = Split(L[i], " \t\r\n,;()");
      if (u.empty()) continue;
      std::string k = lower(u[0]);
      if (k == "input" || k == "output" || k == "wire" || k == "reg" || k == "inout") {
        std::vector<std::string> names;
        if (u[1].find('[') != std::string::npos) names.push_back(u[2]);
        else names.assign(u.begin() + 1, u.end());
        for (auto& n : names) {
          m.port_type[n] = k;
          if (k == "input" || k == "output" || k == "inout") m.in_out_ports.push_back(n);
        }
      } else if (k == "and" || k == "or" || k == "xor" || k == "xnor" || k == "nor" || k == "nand" ||
                 k == "buf" || k == "not" || k == "udp_xbuf" || k == "udp_mux2") {
        Function f;
        f.name = k;
        size_t o = (k == "udp_xbuf" || k == "udp_mux2") ? 2 : 1;
        f.output = u[o];
        for (size_t q = o + 1; q < u.size(); ++q) {
          if (u[q] == "1'b1" || u[q] == "1'b0") {
            std::string nm = "zc_net" + std::to_string(zc++);
            (u[q] == "1'b1" ? m.supply1 : m.supply0).push_back(nm);
            f.input_pins.push_back(nm);
          } else {
            f.input_pins.push_back(u[q]);
          }
        }
        m.function.push_back(f);
      } else if (k == "supply0") {
        for (size_t q = 0; q < u.size(); ++q) m.supply0.push_back(u[q]);
      } else if (k == "supply1") {
        for (size_t q = 0; q < u.size(); ++q) m.supply1.push_back(u[q]);
      } else if (k == "udp_dff" || k == "udp_tlat" || k == "always") {
        ignored = true;
        break;
      } else if (k == "endmodule") {
        break;
      }
    }
    if (ignored) continue;
    D.type_by_name[m.name] = (int)D.types.size();
    D.types.push_back(m);
  }
}

// parse_gv, src/PreProcess.cpp:650-1274 (scalar named connections,
// 1'b0/1'b1 ties, assign)
static void parse_gv(const std::string& path, Design& D) {
  auto L = read_lines(path);
  std::string full;
  bool in_mod = false, header = false;
  int gvn = 0;
  for (auto& raw : L) {
    std::string line = trim(raw);
    if (line.empty()) continue;
    auto t = Split(line, " \t\r\n");
    if (t.empty()) continue;
    if (lower(t[0]) == "module") { in_mod = true; header = line.back() != ';'; continue; }
    if (lower(t[0]) == "endmodule") break;
    if (!in_mod) continue;
    if (header) { if (line.back() == ';') header = false; continue; }
    full += " " + line;
    if (line.back() != ';') continue;
    auto tok = Split(full, " \t\r\n");
    full.clear();
    std::string k = tok[0];
    if (k == "input" || k == "output" || k == "wire" || k == "inout") {
      for (size_t q = 1; q < tok.size(); ++q) {
        std::string n = tok[q];
        if (!n.empty() && (n.back() == ',' || n.back() == ';')) n.pop_back();
        if (!n.empty()) D.net(n);
      }
      continue;
    }
    if (k == "assign") {
      std::string body;
      for (size_t q = 1; q < tok.size(); ++q) body += tok[q];
      if (body.back() == ';') body.pop_back();
      std::string l = body.substr(0, body.find('=')), r = body.substr(body.find('=') + 1);
      D.net(l);
      if (r == "1'b1" || r == "1'b0") D.pinbitValues[l] = TimedValues{{0, r == "1'b1" ? VCD_1 : VCD_0}};
      else { D.net(r); D.assign_pairs[l] = r; }
      continue;
    }
    auto ti = D.type_by_name.find(k);
    if (ti == D.type_by_name.end()) continue;
    const ModuleType& m = D.types[ti->second];
    std::string tl;
    for (size_t q = 1; q < tok.size(); ++q) tl += tok[q];
    auto parts = Split(tl, ".");
    Instance inst;
    inst.name = parts[0].substr(0, parts[0].find('('));
    inst.type = ti->second;
    std::vector<std::string> left = m.in_out_ports;
    for (size_t q = 1; q < parts.size(); ++q) {
      std::string s = parts[q];
      std::string pin = s.substr(0, s.find('('));
      std::string net = s.substr(s.find('(') + 1);
      net = net.substr(0, net.find(')'));
      left.erase(std::find(left.begin(), left.end(), pin));
      bool is_in = m.port_type.at(pin) == "input";
      if (net == "1'b1" || net == "1'b0") {
        std::string nm = "gv_net" + std::to_string(gvn++);
        D.pinbitValues[nm] = TimedValues{{0, net == "1'b1" ? VCD_1 : VCD_0}};
        net = nm;
      }
      D.net(net);
      if (is_in) { inst.in_port_net[pin] = net; inst.in_net.push_back(net); }
      else { inst.out_port_net[pin] = net; inst.out_net.push_back(net); }
    }
    for (auto& p : left)
      if (m.port_type.at(p) == "output") inst.unknown_out.push_back(p);
    // function pin replacement, src/PreProcess.cpp:1192-1265
    for (auto f : m.function) {
      for (auto& p : f.input_pins) {
        auto it = inst.in_port_net.find(p);
        if (it != inst.in_port_net.end()) p = it->second;
      }
      auto ot = inst.out_port_net.find(f.output);
      if (ot != inst.out_port_net.end()) f.output = ot->second;
      inst.functions.push_back(f);
    }
    D.insts.push_back(inst);
  }
}

// parse_sdf, src/PreProcess.cpp:1276-1464
static void parse_sdf(const std::string& path, Design& D) {
  auto L = read_lines(path);
  std::map<std::string, int> by;
  for (size_t i = 0; i < D.insts.size(); ++i) by[D.insts[i].name] = (int)i;
  double factor = 1e3;
  const char* d = " \t\r\n\"'()";
  for (size_t i = 0; i < L.size(); ++i) {
    std::string line = trim(L[i]);
    if (line.empty()) continue;
    auto t = Split(line, d);
    if (t.empty()) continue;
    std::string k = upper(t[0]);
    if (k == "TIMESCALE") {
      if (t[1] == "1ns") factor = 1e3;
      else if (t[1] == "1ps") factor = 1.0;
      else if (t[1] == "1us") factor = 1e6;
      else if (t[1] == "1ms") factor = 1e9;
      else if (t[1] == "1s") factor = 1e12;
      continue;
    }
    if (k != "INSTANCE" || t.size() < 2 || !by.count(t[1])) continue;
    Instance& inst = D.insts[by[t[1]]];
    for (++i; i < L.size(); ++i) {
      std::string l2 = trim(L[i]);
      if (l2.empty()) continue;
      if (l2 == ")") break;
      auto u = Split(l2, d);
      if (u.empty() || upper(u[0]) != "IOPATH") continue;
      Delay dl;
      size_t ip, op, vp;
      std::string e = upper(u[1]);
      if (e == "POSEDGE" || e == "NEGEDGE") { dl.edge = e == "POSEDGE" ? 0 : 1; ip = 2; op = 3; vp = 4; }
      else { dl.edge = 2; ip = 1; op = 2; vp = 3; }
      auto ii = inst.in_port_net.find(u[ip]);
      dl.in_bit = ii == inst.in_port_net.end() ? -1 : D.net(ii->second);
      auto oo = inst.out_port_net.find(u[op]);
      dl.out_bit = oo != inst.out_port_net.end() ? oo->second : u[op];
      auto val = [&](const std::string& s) {
        auto v = Split(s, ":");
        std::stringstream ss(v[1]);
        double x;
        ss >> x;
        return x * factor;
      };
      dl.rise_val = val(u[vp]);
      dl.fall_val = u.size() == vp + 1 ? dl.rise_val : val(u[vp + 1]);
      inst.delays.push_back(dl);
    }
  }
}

struct Vcd {
  TimeUnit unit = TIME_PS;
  unsigned res = 1;
  std::string root = "top";
  std::vector<std::string> names;
  std::vector<TimedValues> waves;
};

static VCDBit code(char c) {
  switch (c) {
    case '0': return VCD_0;
    case '1': return VCD_1;
    case 'x': case 'X': return VCD_X;
    default: return VCD_Z;
  }
}

// VCDParser::init/parse_core/core, src/VCDParser.cpp:194-644 (scalar and
// vector vars; bus bits de-duplicated starting from Z, scalars kept as is)
static void parse_vcd(const std::string& path, Vcd& V) {
  std::ifstream f(path);
  if (!f) { std::fprintf(stderr, "oracle: cannot open %s\n", path.c_str()); std::exit(3); }
  std::string w;
  std::map<std::string, std::vector<int>> ids;
  std::vector<int> vec_bit;
  std::vector<VCDBit> last;
  bool scoped = false;
  while (f >> w) {
    if (w == "$enddefinitions") { while (f >> w && w != "$end") {} break; }
    if (w == "$timescale") {
      std::string acc;
      while (f >> w && w != "$end") acc += w;
      size_t q = 0;
      while (q < acc.size() && isdigit((unsigned char)acc[q])) ++q;
      V.res = q ? (unsigned)std::stoul(acc.substr(0, q)) : 1;
      std::string u = acc.substr(q);
      V.unit = u == "s" ? TIME_S : u == "ms" ? TIME_MS : u == "us" ? TIME_US : u == "ns" ? TIME_NS : TIME_PS;
      continue;
    }
    if (w == "$scope") {
      std::string a, b;
      f >> a >> b;
      if (!scoped) { V.root = b; scoped = true; }
      while (f >> w && w != "$end") {}
      continue;
    }
    if (w == "$var") {
      std::vector<std::string> v;
      while (f >> w && w != "$end") v.push_back(w);
      int size = std::stoi(v[1]);
      std::string id = v[2], ref = v[3];
      std::vector<int> sig;
      if (size > 1 || v.size() > 4) {
        int l = size - 1, r = 0;
        if (v.size() > 4) {
          std::string in = v[4].substr(1, v[4].size() - 2);
          size_t c = in.find(':');
          if (c == std::string::npos) { l = r = std::stoi(in); }
          else { l = std::stoi(in.substr(0, c)); r = std::stoi(in.substr(c + 1)); }
        }
        if (size == 1 && v.size() > 4) {
          sig.push_back((int)V.names.size());
          V.names.push_back(ref + "[" + std::to_string(l) + "]");
          V.waves.emplace_back(); vec_bit.push_back(0); last.push_back(VCD_Z);
        } else {
          int step = l > r ? -1 : 1;
          for (int i = l;; i += step) {
            sig.push_back((int)V.names.size());
            V.names.push_back(ref + "[" + std::to_string(i) + "]");
            V.waves.emplace_back(); vec_bit.push_back(1); last.push_back(VCD_Z);
            if (i == r) break;
          }
        }
      } else {
        sig.push_back((int)V.names.size());
        V.names.push_back(ref);
        V.waves.emplace_back(); vec_bit.push_back(0); last.push_back(VCD_Z);
      }
      auto& dst = ids[id];
      dst.insert(dst.end(), sig.begin(), sig.end());
      continue;
    }
  }
  uint64_t t = 0;
  auto put = [&](int s, VCDBit b) {
    if (vec_bit[s]) { if (b == last[s]) return; last[s] = b; }
    TimedValues& q = V.waves[s];
    if (!q.empty() && q.back().t == t) { q.back().value = b; return; }
    q.push_back({t, b});
  };
  while (f >> w) {
    if (w[0] == '#') { t = std::stoull(w.substr(1)); continue; }
    if (w[0] == '$') continue;
    if (w[0] == 'b' || w[0] == 'B') {
      std::string id;
      f >> id;
      auto& s = ids[id];
      std::string bits = w.substr(1);
      if (bits.size() < s.size()) {
        char pad = (bits[0] == 'x' || bits[0] == 'X' || bits[0] == 'z' || bits[0] == 'Z') ? bits[0] : '0';
        bits = std::string(s.size() - bits.size(), pad) + bits;
      }
      for (size_t q = 0; q < s.size(); ++q) put(s[q], code(bits[bits.size() - s.size() + q]));
      continue;
    }
    if (w[0] == 'r' || w[0] == 'R') { f >> w; continue; }
    auto it = ids.find(w.substr(1));
    if (it == ids.end()) continue;
    for (int s : it->second) put(s, code(w[0]));
  }
}

// calc, src/simulate.cpp:382-870, read with its intended encoding
// (0 -> 0, 1 -> 1, X -> 2, Z -> 3).
static int to_calc(VCDBit b) { return b == VCD_0 ? 0 : b == VCD_1 ? 1 : b == VCD_X ? 2 : 3; }
static VCDBit from_calc(int c) { return c == 0 ? VCD_0 : c == 1 ? VCD_1 : c == 2 ? VCD_X : VCD_Z; }

static VCDBit calc(const std::string& name, const std::vector<VCDBit>& in_b) {
  const int X = 2, Z = 3;
  std::vector<int> in(in_b.size());
  for (size_t i = 0; i < in.size(); ++i) in[i] = to_calc(in_b[i]);
  auto inv = [&](int t) { return (t == 0 || t == 1) ? (1 ^ t) : X; };
  if (name == "and" || name == "nand") {
    int tmp;
    if (in[0] != X && in[1] != X && in[0] != Z && in[1] != Z) tmp = in[0] & in[1];
    else if (in[0] == X || in[0] == Z) tmp = in[1] == 0 ? 0 : X;
    else tmp = in[0] == 0 ? 0 : X;
    for (size_t i = 2; i < in.size(); ++i) {
      switch (tmp) {
        case 0: tmp = 0; break;
        case 1: tmp = in[i] == Z ? X : in[i]; break;
        case 2: case 3: tmp = in[i] == 0 ? 0 : X; break;
        default: tmp = X;
      }
    }
    return from_calc(name == "nand" ? inv(tmp) : tmp);
  }
  if (name == "or" || name == "nor") {
    int tmp;
    if (in[0] != X && in[1] != X && in[0] != Z && in[1] != Z) tmp = in[0] | in[1];
    else if (in[0] == X || in[0] == Z) tmp = in[1] == 1 ? 1 : X;
    else tmp = in[0] == 1 ? 1 : X;
    for (size_t i = 2; i < in.size(); ++i) {
      switch (tmp) {
        case 0: tmp = in[i] == Z ? X : in[i]; break;
        case 1: tmp = 1; break;
        case 2: case 3: tmp = in[i] == 1 ? 1 : X; break;
        default: tmp = X;
      }
    }
    return from_calc(name == "nor" ? inv(tmp) : tmp);
  }
  if (name == "xor" || name == "xnor") {
    int tmp;
    if (in[0] != X && in[1] != X && in[0] != Z && in[1] != Z) tmp = in[0] ^ in[1];
    else if (in[0] == X || in[1] == X) tmp = X;
    else if (in[0] == Z) tmp = in[1] == Z ? 0 : X;
    else tmp = in[0] == Z ? 0 : X;
    for (size_t i = 2; i < in.size(); ++i) {
      switch (tmp) {
        case 0: tmp = in[i] == Z ? X : in[i]; break;
        case 1: tmp = (in[i] == 0 || in[i] == 1) ? (tmp ^ in[i]) : X; break;
        case 2: tmp = X; break;
        case 3: tmp = in[i] == Z ? 0 : X; break;
        default: tmp = X;
      }
    }
    return from_calc(name == "xnor" ? inv(tmp) : tmp);
  }
  if (name == "buf") return from_calc(in[0] == Z ? X : in[0]);
  if (name == "not") return from_calc(inv(in[0]));
  if (name == "udp_xbuf") return from_calc(in[1] == 1 ? (in[0] == X ? 1 : in[0]) : X);
  if (name == "udp_mux2") {
    if (in[2] == 0) return from_calc(in[0]);
    if (in[2] == 1) return from_calc(in[1]);
    if (in[2] == X) return from_calc(in[0] == in[1] ? in[0] : X);
    return VCD_X;
  }
  std::fprintf(stderr, "oracle: function %s is not supported\n", name.c_str());
  std::exit(3);
}

// CheckDelay, src/simulate.cpp:1178-1201
static int CheckDelay(VCDBit pre, VCDBit cur) {
  if ((pre == VCD_0 && cur != VCD_0) || (cur == VCD_1 && pre != VCD_1)) return 1;
  if ((pre == VCD_1 && cur != VCD_1) || (cur == VCD_0 && pre != VCD_0)) return 2;
  if (pre == VCD_X && cur == VCD_Z) return 3;
  if (pre == VCD_Z && cur == VCD_X) return 4;
  return 0;
}

// GetValueAt, src/simulate.cpp:1459-1506
static void GetValueAt(const TimedValues& vals, uint64_t time, VCDBit& cur, VCDBit& last) {
  cur = VCD_X;
  last = VCD_X;
  if (vals.empty()) return;
  auto beg = vals.begin(), end = vals.end();
  auto mid = beg + (end - beg) / 2;
  while (mid != end) {
    if (mid->t < time) {
      cur = mid->value;
      last = mid->value;
      beg = mid + 1;
      mid = beg + (end - beg) / 2;
    } else if (mid->t > time) {
      end = mid;
      mid = beg + (end - beg) / 2;
    } else {
      cur = mid->value;
      if (mid != vals.begin()) last = (mid - 1)->value;
      break;
    }
  }
}

// Process, src/simulate.cpp:1365-1457
static void Process(const std::map<int, const TimedValues*>& inValues, std::deque<TimedValueSim>& q) {
  bool first = true;
  for (auto& kv : inValues) {
    int name = kv.first;
    const TimedValues* vals = kv.second;
    if (first) {
      for (auto& tv : *vals) {
        TimedValueSim s;
        s.time = tv.t;
        s.values[name] = tv.value;
        q.push_back(s);
      }
      first = false;
      continue;
    }
    for (auto& tv : *vals) {
      auto beg = q.begin(), end = q.end();
      auto mid = beg + (end - beg) / 2;
      bool found = false;
      while (mid != end) {
        if (mid->time < tv.t) { beg = mid + 1; mid = beg + (end - beg) / 2; }
        else if (mid->time > tv.t) { end = mid; mid = beg + (end - beg) / 2; }
        else { found = true; break; }
      }
      if (found) { mid->values[name] = tv.value; continue; }
      TimedValueSim s;
      s.time = tv.t;
      s.values[name] = tv.value;
      q.insert(mid, s);
    }
  }
}

// RunFunction, src/simulate.cpp:1203-1363
static VCDBit RunFunction(const Design& D, const Instance& inst, const std::string& out,
                          const std::map<std::string, VCDBit>& cur_in) {
  const ModuleType& m = D.types[inst.type];
  std::map<std::string, VCDBit> inter;
  for (const Function& f : inst.functions) {
    std::vector<VCDBit> data;
    for (auto& p : f.input_pins) {
      auto a = cur_in.find(p);
      if (a != cur_in.end()) { data.push_back(a->second); continue; }
      auto b = inter.find(p);
      if (b != inter.end()) { data.push_back(b->second); continue; }
      if (std::find(m.supply1.begin(), m.supply1.end(), p) != m.supply1.end()) { data.push_back(VCD_1); continue; }
      if (std::find(m.supply0.begin(), m.supply0.end(), p) != m.supply0.end()) { data.push_back(VCD_0); continue; }
      data.push_back(VCD_X);
    }
    VCDBit r = calc(f.name, data);
    if (f.output == out) return r;
    inter[f.output] = r;
  }
  std::fprintf(stderr, "oracle: no function drives %s of %s\n", out.c_str(), inst.name.c_str());
  std::exit(3);
}

// AddDelay, src/simulate.cpp:1012-1176 (asserts compiled out)
static uint64_t AddDelay(const std::vector<const Delay*>& rel, unsigned num_in, int s, uint64_t t,
                         const std::map<int, VCDBit>& cur_in, const std::map<int, VCDBit>& last_in,
                         TimeUnit unit) {
  auto pick = [&](const Delay* d, int sense) {
    switch (sense) {
      case 1: return d->rise_val;
      case 2: return d->fall_val;
      case 3: return std::max(d->rise_val, d->fall_val);
      case 4: return std::min(d->rise_val, d->fall_val);
      default: return 1.0;
    }
  };
  auto in_edge = [&](int net) {
    int in_sense = CheckDelay(last_in.at(net), cur_in.at(net));
    if (in_sense == 1) return 0;
    if (in_sense == 2) return 1;
    return 0;  // "regarded as posedge"
  };
  double delay_val = 1.0, factor = 1.0;
  if (rel.empty()) delay_val = 1.0;
  else if (num_in == 1) {
    if (rel.size() == 1) delay_val = pick(rel[0], s);
    else {
      int e = in_edge(rel[0]->in_bit);
      for (const Delay* d : rel)
        if (d->edge == e) { delay_val = pick(d, s); break; }
    }
  } else {
    double mn = std::numeric_limits<double>::max();
    for (const Delay* d : rel) {
      if (d->edge == 2) mn = std::min(pick(d, s), mn);
      else if (in_edge(d->in_bit) == d->edge) mn = std::min(pick(d, s), mn);
    }
    delay_val = mn == std::numeric_limits<double>::max() ? 1.0 : mn;
  }
  if (unit == TIME_S) factor = 1e-12;
  else if (unit == TIME_MS) factor = 1e-9;
  else if (unit == TIME_US) factor = 1e-6;
  else if (unit == TIME_NS) factor = 1e-3;
  else factor = 1.0;
  return t + static_cast<unsigned long long>(delay_val * factor);
}

struct Result {
  std::map<std::string, TimedValues> out_value;  // gate outputs (incl. inst[pin])
  uint64_t n_events = 0;
};

// SimulateInst + Calculate, src/simulate.cpp:872-1011
static void SimulateInst(const Design& D, const Instance& inst,
                         const std::map<int, const TimedValues*>& bitValues,
                         const std::vector<int>& in_ids, TimeUnit unit, Result& R) {
  std::vector<std::pair<std::string, std::string>> outs;  // (_out, _out_name)
  for (auto& o : inst.out_net) outs.push_back({o, o});
  for (auto& o : inst.unknown_out) outs.push_back({o, inst.name + "[" + o + "]"});
  for (auto& oo : outs) {
    const std::string& _out = oo.first;
    TimedValues& q = R.out_value[oo.second];
    q.clear();
    q.push_back({0, VCD_X});
    std::deque<TimedValueSim> in_tv;
    Process(bitValues, in_tv);
    VCDBit pre_out = VCD_X;
    for (auto& temp : in_tv) {
      R.n_events++;
      uint64_t cur_time = temp.time;
      std::vector<const Delay*> rel;
      for (auto& d : inst.delays)
        if (d.out_bit == _out && temp.values.count(d.in_bit)) rel.push_back(&d);
      std::map<int, VCDBit> cur_in, last_in;
      std::map<std::string, VCDBit> cur_by_name;
      for (int id : in_ids) {
        VCDBit c, l;
        GetValueAt(*bitValues.at(id), cur_time, c, l);
        cur_in[id] = c;
        last_in[id] = l;
        cur_by_name[D.net_name[id]] = c;
      }
      VCDBit cur_out = RunFunction(D, inst, _out, cur_by_name);
      int out_sense = CheckDelay(pre_out, cur_out);
      if (out_sense == 0) continue;
      uint64_t out_time = AddDelay(rel, (unsigned)temp.values.size(), out_sense, cur_time, cur_in, last_in, unit);
      TimedValue last_tv = q.back();
      while (last_tv.t >= out_time) {
        if (q.size() <= 1) { q.pop_back(); break; }
        q.pop_back();
        last_tv = q.back();
      }
      if (last_tv.value != cur_out) q.push_back({out_time, cur_out});
      pre_out = q.back().value;
    }
  }
}

static std::vector<std::vector<int>> levelize(const Design& D) {
  std::map<std::string, int> driver;
  for (size_t i = 0; i < D.insts.size(); ++i)
    for (auto& o : D.insts[i].out_net) driver[o] = (int)i;
  std::vector<std::vector<int>> rd(D.insts.size());
  std::vector<int> indeg(D.insts.size(), 0);
  for (size_t i = 0; i < D.insts.size(); ++i) {
    std::vector<int> srcs;
    for (auto& n : D.insts[i].in_net) {
      std::string nn = n;
      auto a = D.assign_pairs.find(n);
      auto it = driver.find(nn);
      if (it == driver.end() && a != D.assign_pairs.end()) it = driver.find(a->second);
      if (it != driver.end()) srcs.push_back(it->second);
    }
    std::sort(srcs.begin(), srcs.end());
    srcs.erase(std::unique(srcs.begin(), srcs.end()), srcs.end());
    for (int s : srcs) { rd[s].push_back((int)i); indeg[i]++; }
  }
  std::vector<std::vector<int>> levels;
  std::vector<int> fr;
  for (size_t i = 0; i < D.insts.size(); ++i) if (!indeg[i]) fr.push_back((int)i);
  size_t placed = 0;
  while (!fr.empty()) {
    levels.push_back(fr);
    placed += fr.size();
    std::vector<int> nx;
    for (int g : fr) for (int r : rd[g]) if (--indeg[r] == 0) nx.push_back(r);
    std::sort(nx.begin(), nx.end());
    fr.swap(nx);
  }
  if (placed != D.insts.size()) { std::fprintf(stderr, "oracle: combinational loop\n"); std::exit(3); }
  return levels;
}

// DumpSaif loop, src/simulate.cpp:102-163
static void saif_of(const TimedValues& tvs, uint64_t start, uint64_t end, uint64_t v[3]) {
  uint64_t t1 = 0, t0 = 0, tx = 0, tz = 0, last_time = start;
  VCDBit last_state = VCD_X;
  for (auto& tv : tvs) {
    if (tv.t <= start) { last_state = tv.value; continue; }
    if (tv.t > end) break;
    uint64_t d = tv.t - last_time;
    if (last_state == VCD_X) tx += d;
    else if (last_state == VCD_0) t0 += d;
    else if (last_state == VCD_1) t1 += d;
    else tz += d;
    last_time = tv.t;
    last_state = tv.value;
  }
  if (last_time < end) {
    uint64_t d = end - last_time;
    if (last_state == VCD_X) tx += d;
    else if (last_state == VCD_0) t0 += d;
    else if (last_state == VCD_1) t1 += d;
    else tz += d;
  }
  v[0] = t0; v[1] = t1; v[2] = tx + tz;
}

struct Output {
  std::string saif;
  uint64_t waves_hash = 0;
  uint64_t n_events = 0;
  double sim_seconds = 0;
  Result res;
};

static uint64_t wave_hash(const std::string& name, const TimedValues& q) {
  uint64_t h = 1469598103934665603ull;
  for (unsigned char c : name) h = (h ^ c) * 1099511628211ull;
  for (auto& tv : q) {
    uint32_t w = (uint32_t)(tv.t << 2) | (uint32_t)tv.value;
    for (int b = 0; b < 4; ++b) h = (h ^ ((w >> (8 * b)) & 0xFF)) * 1099511628211ull;
  }
  uint64_t z = h + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// main's level loop, src/simulate.cpp:1569-1668, then DumpSaif
static void run(const char* vlib, const char* gv, const char* sdf, const char* vcd, uint64_t start,
                uint64_t end, Output& O, uint64_t max_events = 0) {
  Design D;
  parse_vlib(vlib, D);
  parse_gv(gv, D);
  parse_sdf(sdf, D);
  Vcd V;
  parse_vcd(vcd, V);
  std::map<std::string, int> vcd_idx;
  for (size_t i = 0; i < V.names.size(); ++i) vcd_idx[V.names[i]] = (int)i;
  auto levels = levelize(D);
  auto t0 = std::chrono::steady_clock::now();
  Result& R = O.res;
  for (auto& lv : levels) {
    for (int ii : lv) {
      const Instance& inst = D.insts[ii];
      std::map<int, const TimedValues*> bitValues;
      std::vector<int> ids;
      std::vector<TimedValues> owned;
      owned.reserve(inst.in_net.size());
      for (auto& in : inst.in_net) {
        int id = D.net_id.at(in);
        const TimedValues* src = nullptr;
        auto resolve = [&](const std::string& n) -> const TimedValues* {
          auto a = vcd_idx.find(n);
          if (a != vcd_idx.end()) return &V.waves[a->second];
          auto b = R.out_value.find(n);
          if (b != R.out_value.end()) return &b->second;
          auto c = D.pinbitValues.find(n);
          if (c != D.pinbitValues.end()) return &c->second;
          return nullptr;
        };
        src = resolve(in);
        if (!src) {
          auto a = D.assign_pairs.find(in);
          if (a != D.assign_pairs.end()) src = resolve(a->second);
        }
        if (!src) { std::fprintf(stderr, "oracle: net %s has no driver\n", in.c_str()); std::exit(3); }
        if (!bitValues.count(id)) ids.push_back(id);
        bitValues[id] = src;
      }
      SimulateInst(D, inst, bitValues, ids, V.unit, R);
      if (max_events && R.n_events >= max_events) break;
    }
    if (max_events && R.n_events >= max_events) break;
  }
  auto t1 = std::chrono::steady_clock::now();
  O.sim_seconds = std::chrono::duration<double>(t1 - t0).count();
  O.n_events = R.n_events;
  // SAIF: gate outputs, visible constants, every VCD signal; sorted by name
  std::map<std::string, std::array<uint64_t, 3>> recs;
  for (auto& kv : R.out_value) { uint64_t v[3]; saif_of(kv.second, start, end, v); recs[kv.first] = {v[0], v[1], v[2]}; }
  for (auto& kv : D.pinbitValues) {
    if (kv.first.substr(0, 6) == "gv_net") continue;
    uint64_t v[3]; saif_of(kv.second, start, end, v); recs[kv.first] = {v[0], v[1], v[2]};
  }
  for (size_t i = 0; i < V.names.size(); ++i) {
    uint64_t v[3]; saif_of(V.waves[i], start, end, v); recs[V.names[i]] = {v[0], v[1], v[2]};
  }
  const char* units[] = {"s", "ms", "us", "ns", "ps"};
  std::string o;
  char buf[200];
  std::snprintf(buf, sizeof buf, "(TIMESCALE %u %s)\n(DURATION %llu)\n(INSTANCE root\n   (NET\n", V.res,
                units[V.unit], (unsigned long long)(end - start));
  o += buf;
  std::vector<std::pair<std::string, std::array<uint64_t, 3>>> hier;
  for (auto& kv : recs) {
    const std::string& n = kv.first;
    size_t lb = n.find('[');
    if (lb != std::string::npos && lb + 1 < n.size() && !(n[lb + 1] >= '0' && n[lb + 1] <= '9')) {
      hier.push_back(kv);
      continue;
    }
    std::snprintf(buf, sizeof buf, "         (T0 %llu) (T1 %llu) (TX %llu)\n      )\n",
                  (unsigned long long)kv.second[0], (unsigned long long)kv.second[1],
                  (unsigned long long)kv.second[2]);
    o += "      (" + n + "\n" + buf;
  }
  o += "   )\n\n";
  for (auto& kv : hier) {
    const std::string& n = kv.first;
    size_t lb = n.find('['), rb = n.find(']');
    o += "   (INSTANCE " + n.substr(0, lb) + "\n      (NET\n         (" + n.substr(lb + 1, rb - lb - 1) + "\n";
    std::snprintf(buf, sizeof buf, "            (T0 %llu) (T1 %llu) (TX %llu)\n         )\n      )\n   )\n",
                  (unsigned long long)kv.second[0], (unsigned long long)kv.second[1],
                  (unsigned long long)kv.second[2]);
    o += buf;
  }
  o += ")\n";
  O.saif = o;
  uint64_t h = 0;
  for (auto& kv : R.out_value) h += wave_hash(kv.first, kv.second);
  O.waves_hash = h;
}

}  // namespace oracle

// ------------------------------------------------------------------ C API ---
extern "C" {

// Runs the oracle; writes the SAIF to saif_path (if non-null) and each gate
// output's waveform as "name t:v t:v ...\n" lines to waves_path (if
// non-null).  Returns 0.  max_events > 0 stops after that many gate events
// (CPU-baseline sampling; results are then partial).
int oracle_run(const char* vlib, const char* gv, const char* sdf, const char* vcd,
               unsigned long long start, unsigned long long end, const char* saif_path,
               const char* waves_path, unsigned long long max_events,
               unsigned long long* n_events, unsigned long long* waves_hash, double* sim_seconds) {
  oracle::Output O;
  oracle::run(vlib, gv, sdf, vcd, start, end, O, max_events);
  if (saif_path) {
    std::ofstream f(saif_path, std::ios::binary);
    f << O.saif;
  }
  if (waves_path) {
    std::ofstream f(waves_path, std::ios::binary);
    for (auto& kv : O.res.out_value) {
      f << kv.first;
      for (auto& tv : kv.second) f << ' ' << tv.t << ':' << (int)tv.value;
      f << '\n';
    }
  }
  if (n_events) *n_events = O.n_events;
  if (waves_hash) *waves_hash = O.waves_hash;
  if (sim_seconds) *sim_seconds = O.sim_seconds;
  return 0;
}

// Single primitive evaluation with calc() (codes are VCDBit values).
int oracle_calc(const char* name, const int* in, int n) {
  std::vector<oracle::VCDBit> v;
  for (int i = 0; i < n; ++i) v.push_back((oracle::VCDBit)in[i]);
  return (int)oracle::calc(name, v);
}

int oracle_check_delay(int pre, int cur) {
  return oracle::CheckDelay((oracle::VCDBit)pre, (oracle::VCDBit)cur);
}

}  // extern "C"

#ifdef ORACLE_MAIN
int main(int argc, char** argv) {
  if (argc < 8) {
    std::printf("usage: cpu_ref vlib gv sdf vcd dumpon dumpoff saif_out [max_events]\n");
    return 1;
  }
  unsigned long long nev = 0, h = 0;
  double secs = 0;
  oracle_run(argv[1], argv[2], argv[3], argv[4], std::strtoull(argv[5], nullptr, 10),
             std::strtoull(argv[6], nullptr, 10), argv[7], nullptr,
             argc > 8 ? std::strtoull(argv[8], nullptr, 10) : 0ull, &nev, &h, &secs);
  std::printf("{\"n_events\": %llu, \"sim_seconds\": %.6f, \"events_per_s\": %.1f, \"waves_hash\": \"%016llx\"}\n",
              nev, secs, secs > 0 ? nev / secs : 0.0, h);
  return 0;
}
#endif
