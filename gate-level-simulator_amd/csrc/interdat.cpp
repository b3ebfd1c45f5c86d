#include "interdat.h"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <map>
#include <set>

#include "../../include/gls_status.h"

namespace gls {

namespace {

// ------------------------------------------------- text archive writer ---
// Boost text-archive syntax: space-separated tokens; a string is its length
// then its characters; a collection is its count, an item version (0), then
// its items; a user class is preceded, the first time its type appears, by
// its class information (tracking level 0, version 0).
class ArchiveOut {
 public:
  ArchiveOut() {
    str("serialization::archive");
    num(17);
  }
  void num(long long v) {
    char b[32];
    std::snprintf(b, sizeof b, "%lld", v);
    s_ += b;
    s_ += ' ';
  }
  void real(double v) {
    char b[40];
    std::snprintf(b, sizeof b, "%.17g", v);
    s_ += b;
    s_ += ' ';
  }
  void str(const std::string& v) {
    num((long long)v.size());
    s_ += v;
    s_ += ' ';
  }
  void cls(const char* type) {
    if (seen_.insert(type).second) {
      num(0);
      num(0);
    }
  }
  void count(size_t n) {
    num((long long)n);
    num(0);
  }
  void strs(const std::vector<std::string>& v) {
    count(v.size());
    for (auto& x : v) str(x);
  }
  template <typename T>
  void nums(const std::vector<T>& v) {
    count(v.size());
    for (auto x : v) num((long long)x);
  }
  template <typename T>
  void reals(const std::vector<T>& v) {
    count(v.size());
    for (auto x : v) real((double)x);
  }
  std::string& text() { return s_; }

 private:
  std::string s_;
  std::set<std::string> seen_;
};

const char* kPrimName[] = {"and", "or", "xor", "xnor", "nor", "nand", "buf", "not", "udp_xbuf", "udp_mux2"};

// Pin record of the reference (src/VCDTypes.h:112-138): type, name, size,
// lindex, rindex; single bits here.
void put_pin(ArchiveOut& a, const std::string& name, const char* type) {
  a.cls("Pin");
  a.str(type);
  a.str(name);
  a.num(1);
  a.num(-1);
  a.num(-1);
}

// Function record (src/Function.h:13-66): name, id, input pins, input data
// (none), output pins, output data (X).
void put_function(ArchiveOut& a, const Prim& p, const std::vector<std::string>& ins,
                  const std::string& out) {
  a.cls("Function");
  a.str(kPrimName[p.op]);
  a.num(p.op);
  a.strs(ins);
  a.count(0);
  a.strs({out});
  a.num(1);
}

// ModuleType (src/VCDTypes.h:190-211): name, ports (input/output/wire
// records), port order, supply1 and supply0 nets, function chain.
void put_module(ArchiveOut& a, const CellType& ct) {
  a.cls("ModuleType");
  a.str(ct.name);
  std::map<std::string, const char*> ports;
  for (auto& p : ct.inputs) ports[p] = "input";
  for (auto& p : ct.outputs) ports[p] = "output";
  for (auto& pr : ct.prims)
    if (!ports.count(pr.out)) ports[pr.out] = "wire";
  a.count(ports.size());
  for (auto& kv : ports) {
    a.str(kv.first);
    put_pin(a, kv.first, kv.second);
  }
  a.strs(ct.ports);
  a.strs(ct.supply1);
  a.strs(ct.supply0);
  a.count(ct.prims.size());
  for (auto& pr : ct.prims) put_function(a, pr, pr.ins, pr.out);
}

struct InstView {
  std::vector<std::string> in_net, out_net;        // the reference's in_net / out_net
  std::map<std::string, std::string> in_port_net, out_port_net;
  std::vector<std::vector<int>> fid;               // function_id_vec
  std::vector<std::vector<std::string>> fins;      // function input pins after renaming
  std::vector<std::string> fout;
};

// The reference's per-instance view (src/PreProcess.cpp:865-1268): input and
// output nets in statement order, port maps, and the function chain with
// pins renamed to nets and numbered (inputs by position in in_net, outputs
// after them, internal wires from in + out + unconnected outputs, supply1 17,
// supply0 16).
InstView inst_view(const Instance& in, const CellType& ct, const Netlist& nl) {
  InstView v;
  for (auto& p : in.in_pins) {
    v.in_net.push_back(nl.net_names[p.second]);
    v.in_port_net[p.first] = nl.net_names[p.second];
  }
  for (auto& p : in.out_pins) {
    v.out_net.push_back(nl.net_names[p.second]);
    v.out_port_net[p.first] = nl.net_names[p.second];
  }
  int pin_id = (int)(v.in_net.size() + v.out_net.size() + in.unknown_out.size());
  std::map<std::string, int> wires;
  auto code = [&](const std::string& pin, bool is_out, std::string& renamed) -> int {
    renamed = pin;
    if (!is_out) {
      auto it = v.in_port_net.find(pin);
      if (it != v.in_port_net.end()) {
        renamed = it->second;
        return (int)(std::find(v.in_net.begin(), v.in_net.end(), it->second) - v.in_net.begin());
      }
    } else {
      auto it = v.out_port_net.find(pin);
      if (it != v.out_port_net.end()) {
        renamed = it->second;
        return (int)(std::find(v.out_net.begin(), v.out_net.end(), it->second) - v.out_net.begin() +
                     v.in_net.size());
      }
    }
    auto w = wires.find(pin);
    if (w != wires.end()) return w->second;
    if (std::find(ct.supply1.begin(), ct.supply1.end(), pin) != ct.supply1.end()) return 17;
    if (std::find(ct.supply0.begin(), ct.supply0.end(), pin) != ct.supply0.end()) return 16;
    wires[pin] = pin_id;
    return pin_id++;
  };
  for (auto& pr : ct.prims) {
    std::vector<int> row{pr.op};
    std::vector<std::string> ins;
    std::string nm;
    for (auto& p : pr.ins) {
      row.push_back(code(p, false, nm));
      ins.push_back(nm);
    }
    row.push_back(code(pr.out, true, nm));
    v.fid.push_back(row);
    v.fins.push_back(ins);
    v.fout.push_back(nm);
  }
  return v;
}

}  // namespace

std::string inter_archive_text(const Library& lib, const Netlist& nl, const Program& pg) {
  const size_t NI = nl.insts.size();
  std::vector<InstView> views;
  views.reserve(NI);
  for (auto& in : nl.insts) views.push_back(inst_view(in, lib.types[in.type], nl));
  // drivers, readers, levels (an instance sits at the level of its gates;
  // one whose outputs the stimulus overrides has no gate and sits at 0)
  std::map<std::string, int> out_net_from_id;
  std::map<std::string, std::vector<int>> net_instance_map;
  for (size_t i = 0; i < NI; ++i) {
    for (auto& n : views[i].out_net) out_net_from_id[n] = (int)i;
    for (auto& n : views[i].in_net) net_instance_map[n].push_back((int)i);
  }
  std::vector<int> inst_level(NI, -1);
  for (size_t l = 0; l + 1 < pg.level_start.size(); ++l)
    for (uint32_t g = pg.level_start[l]; g < pg.level_start[l + 1]; ++g)
      if (inst_level[pg.gates[g].inst] < 0) inst_level[pg.gates[g].inst] = (int)l;
  size_t nlev = std::max<size_t>(1, pg.level_start.size() - 1);
  std::vector<std::vector<int>> levels(nlev);
  for (size_t i = 0; i < NI; ++i) levels[std::max(0, inst_level[i])].push_back((int)i);
  std::vector<int> pos_at_level(NI, 0);
  for (auto& L : levels)
    for (size_t j = 0; j < L.size(); ++j) pos_at_level[L[j]] = (int)j;
  // provenance of every instance input (src/process.cpp:276-372)
  std::map<std::string, int> initial_net_map;
  struct From { std::vector<int> id, level, pos, order; std::vector<std::string> info; };
  std::vector<From> from(NI);
  auto add_from = [&](size_t i, const std::string& net, int depth) {
    auto d = out_net_from_id.find(net);
    if (d != out_net_from_id.end()) {
      const int id = d->second;
      const auto& on = views[id].out_net;
      const int order = on.size() <= 1 ? 0 : (int)(std::find(on.begin(), on.end(), net) - on.begin());
      from[i].id.push_back(id);
      from[i].info.push_back(nl.insts[id].name);
      from[i].level.push_back(std::max(0, inst_level[id]));
      from[i].pos.push_back(pos_at_level[id]);
      from[i].order.push_back(order);
      return true;
    }
    const int n = nl.net(net);
    if (n >= 0 && nl.const_val[n] >= 0) {
      from[i].id.push_back(-2);
      from[i].info.push_back(net);
      from[i].level.push_back(-2);
      from[i].pos.push_back(nl.const_val[n]);
      from[i].order.push_back(0);
      return true;
    }
    if (depth == 0 && n >= 0 && nl.alias_of[n] >= 0) return false;  // resolved by the caller
    if (!initial_net_map.count(net)) {
      const int id = (int)initial_net_map.size();
      initial_net_map[net] = id;
    }
    from[i].id.push_back(-1);
    from[i].info.push_back(net);
    from[i].level.push_back(-1);
    from[i].pos.push_back(initial_net_map[net]);
    from[i].order.push_back(0);
    return true;
  };
  for (size_t i = 0; i < NI; ++i)
    for (auto& net : views[i].in_net)
      if (!add_from(i, net, 0)) add_from(i, nl.net_names[nl.alias_of[nl.net(net)]], 1);

  ArchiveOut a;
  a.cls("inter");
  // PreProcess p (src/PreProcess.h:74-90)
  a.cls("PreProcess");
  {
    std::map<std::string, const CellType*> types;
    for (auto& t : lib.types) types[t.name] = &t;
    a.count(types.size());
    for (auto& kv : types) {
      a.str(kv.first);
      put_module(a, *kv.second);
    }
  }
  a.strs(lib.ignored);
  a.str(nl.top);
  {
    std::map<std::string, size_t> by_name;
    for (size_t i = 0; i < NI; ++i) by_name[nl.insts[i].name] = i;
    a.count(by_name.size());
    for (auto& kv : by_name) {
      const size_t i = kv.second;
      const Instance& in = nl.insts[i];
      const CellType& ct = lib.types[in.type];
      const InstView& v = views[i];
      a.str(kv.first);
      // Instance (src/Instance.h:107-128)
      a.cls("Instance");
      put_module(a, ct);
      a.str(in.name);
      a.count(v.in_port_net.size());
      for (auto& p : v.in_port_net) {
        a.str(p.first);
        a.str(p.second);
      }
      a.strs(v.in_net);
      a.count(v.out_port_net.size());
      for (auto& p : v.out_port_net) {
        a.str(p.first);
        a.str(p.second);
      }
      a.strs(v.out_net);
      a.count(0);  // unknown_in
      a.strs(in.unknown_out);
      a.count(ct.prims.size());
      for (size_t q = 0; q < ct.prims.size(); ++q) put_function(a, ct.prims[q], v.fins[q], v.fout[q]);
      a.count(v.fid.size());
      for (auto& row : v.fid) a.nums(row);
      a.count(in.delays.size());
      for (auto& d : in.delays) {
        a.cls("Delay");
        a.num(d.edge);
        a.str(d.in_net >= 0 ? nl.net_names[d.in_net] : std::string());
        a.str(d.out_net >= 0 ? nl.net_names[d.out_net] : in.unknown_out[-2 - d.out_net]);
        a.real(d.rise);
        a.real(d.fall);
      }
      a.nums(from[i].id);
      a.strs(from[i].info);
      a.nums(from[i].level);
      a.nums(from[i].pos);
      a.nums(from[i].order);
    }
  }
  {
    std::vector<std::string> names;
    for (auto& in : nl.insts) names.push_back(in.name);
    a.strs(names);
  }
  a.count(net_instance_map.size());
  for (auto& kv : net_instance_map) {
    a.str(kv.first);
    a.nums(kv.second);
  }
  a.count(out_net_from_id.size());
  for (auto& kv : out_net_from_id) {
    a.str(kv.first);
    a.num(kv.second);
  }
  {
    std::map<std::string, const char*> pins;
    for (size_t n = 0; n < nl.net_names.size(); ++n)
      pins[nl.net_names[n]] = nl.net_is_input[n] ? "input" : "wire";
    a.count(pins.size());
    for (auto& kv : pins) {
      a.str(kv.first);
      put_pin(a, kv.first, kv.second);
    }
    std::vector<std::string> bits;
    for (auto& kv : pins) bits.push_back(kv.first);
    a.strs(bits);
  }
  {
    // constants: a pointer to a one-entry waveform (value from time 0)
    std::map<std::string, int> consts;
    for (size_t n = 0; n < nl.net_names.size(); ++n)
      if (nl.const_val[n] >= 0) consts[nl.net_names[n]] = nl.const_val[n];
    a.count(consts.size());
    for (auto& kv : consts) {
      a.str(kv.first);
      a.num(1);  // pointer: object present
      a.count(1);
      a.cls("TimedValue");
      a.num(0);
      a.num(kv.second);
    }
    std::map<std::string, std::string> assigns;
    for (size_t n = 0; n < nl.net_names.size(); ++n)
      if (nl.alias_of[n] >= 0) assigns[nl.net_names[n]] = nl.net_names[nl.alias_of[n]];
    a.count(assigns.size());
    for (auto& kv : assigns) {
      a.str(kv.first);
      a.str(kv.second);
    }
    a.count(0);  // Result
  }
  // levels, initial nets, then the per-level flat arrays
  // (src/process.cpp:423-566)
  a.count(levels.size());
  for (auto& L : levels) a.nums(L);
  a.count(initial_net_map.size());
  for (auto& kv : initial_net_map) {
    a.str(kv.first);
    a.num(kv.second);
  }
  const size_t NL = levels.size();
  std::vector<unsigned> in_start(NL), out_start(NL), delay_start(NL), delay_width(NL), fn_start(NL),
      fn_width(NL);
  std::vector<short> data_in(NI), data_out(NI), delay_num(NI), fn_func(NI), fn_val(NI);
  std::vector<unsigned> val_start(NI);
  unsigned sum = 0, dstart = 0, fstart = 0;
  for (size_t l = 0; l < NL; ++l) {
    unsigned vw = 0, dw = 0, fw = 0;
    in_start[l] = out_start[l] = sum;
    fn_start[l] = fstart;
    delay_start[l] = dstart;
    for (size_t j = 0; j < levels[l].size(); ++j) {
      const int i = levels[l][j];
      const InstView& v = views[i];
      data_in[j + sum] = (short)v.in_net.size();
      data_out[j + sum] = (short)v.out_net.size();
      val_start[j + sum] = vw;
      vw += (unsigned)v.in_net.size();
      delay_num[j + sum] = (short)nl.insts[i].delays.size();
      dw = std::max<unsigned>(dw, (unsigned)nl.insts[i].delays.size());
      unsigned fwid = 0;
      for (auto& row : v.fid) fwid = std::max<unsigned>(fwid, (unsigned)row.size());
      fw = std::max<unsigned>(fw, fwid * (unsigned)v.fid.size());
      fn_func[j + sum] = (short)v.fid.size();
      fn_val[j + sum] = (short)fwid;
    }
    sum += (unsigned)levels[l].size();
    dstart += (unsigned)levels[l].size() * dw;
    fstart += (unsigned)levels[l].size() * fw;
    delay_width[l] = dw;
    fn_width[l] = fw;
  }
  std::vector<short> d_edge(dstart, -1), d_in(dstart, -1), d_out(dstart, -1), fns(fstart, -1);
  std::vector<float> d_rise(dstart, -1.0f), d_fall(dstart, -1.0f);
  for (size_t l = 0; l < NL; ++l)
    for (size_t j = 0; j < levels[l].size(); ++j) {
      const int i = levels[l][j];
      const InstView& v = views[i];
      const Instance& in = nl.insts[i];
      for (size_t q = 0; q < in.delays.size(); ++q) {
        const size_t at = delay_start[l] + j * delay_width[l] + q;
        const IoPath& d = in.delays[q];
        d_edge[at] = (short)d.edge;
        const std::string inb = d.in_net >= 0 ? nl.net_names[d.in_net] : std::string();
        d_in[at] = (short)(std::find(v.in_net.begin(), v.in_net.end(), inb) - v.in_net.begin());
        const std::string outb = d.out_net >= 0 ? nl.net_names[d.out_net] : std::string();
        d_out[at] = (short)(std::find(v.out_net.begin(), v.out_net.end(), outb) - v.out_net.begin());
        d_rise[at] = (float)d.rise;
        d_fall[at] = (float)d.fall;
      }
      const size_t h = v.fid.size();
      if (!h) continue;
      const size_t w = fn_width[l] / h;
      for (size_t r = 0; r < h; ++r)
        for (size_t c = 0; c < w; ++c)
          fns[fn_start[l] + j * fn_width[l] + r * w + c] = c < v.fid[r].size() ? (short)v.fid[r][c] : -1;
    }
  a.nums(in_start);
  a.nums(out_start);
  a.nums(delay_start);
  a.nums(delay_width);
  a.nums(fn_start);
  a.nums(fn_width);
  a.nums(data_in);
  a.nums(data_out);
  a.nums(val_start);
  a.nums(delay_num);
  a.nums(fn_func);
  a.nums(fn_val);
  a.nums(d_edge);
  a.nums(d_in);
  a.nums(d_out);
  a.reals(d_rise);
  a.reals(d_fall);
  a.nums(fns);
  // OutMaps: per level, output net -> (position at level, output order)
  a.count(NL);
  for (size_t l = 0; l < NL; ++l) {
    std::map<std::string, std::vector<unsigned>> m;
    for (size_t j = 0; j < levels[l].size(); ++j) {
      const auto& on = views[levels[l][j]].out_net;
      for (size_t k = 0; k < on.size(); ++k)
        if (!m.count(on[k])) m[on[k]] = {(unsigned)j, (unsigned)k};
    }
    a.count(m.size());
    for (auto& kv : m) {
      a.str(kv.first);
      a.nums(kv.second);
    }
  }
  return std::move(a.text());
}

// --------------------------------------------------- compiled section ---
namespace {

// The compiled section: magic, its byte length on a line, the binary
// records, and a fixed-length trailer holding the offset of the magic (the
// section is found from the end of the file, whatever the archive holds).
const char kMagic[] = "\nGLS-CSR 1\n";
constexpr size_t kTrailer = 33;  // "\nGLS-CSR-AT " + 20 digits + "\n"

class BinOut {
 public:
  void u64(uint64_t v) { s_.append((const char*)&v, 8); }
  void i64(int64_t v) { u64((uint64_t)v); }
  void f64(double v) { s_.append((const char*)&v, 8); }
  void str(const std::string& v) {
    u64(v.size());
    s_ += v;
  }
  void strs(const std::vector<std::string>& v) {
    u64(v.size());
    for (auto& x : v) str(x);
  }
  template <typename T>
  void pod(const std::vector<T>& v) {
    u64(v.size());
    s_.append((const char*)v.data(), v.size() * sizeof(T));
  }
  std::string& text() { return s_; }

 private:
  std::string s_;
};

class BinIn {
 public:
  BinIn(const std::string& s, size_t p) : s_(s), p_(p) {}
  uint64_t u64() {
    need(8);
    uint64_t v;
    std::memcpy(&v, s_.data() + p_, 8);
    p_ += 8;
    return v;
  }
  int64_t i64() { return (int64_t)u64(); }
  double f64() {
    const uint64_t b = u64();
    double v;
    std::memcpy(&v, &b, 8);
    return v;
  }
  std::string str() {
    const uint64_t n = u64();
    need(n);
    std::string v = s_.substr(p_, n);
    p_ += n;
    return v;
  }
  std::vector<std::string> strs() {
    const uint64_t n = u64();
    if (n > s_.size()) bad();
    std::vector<std::string> v(n);
    for (auto& x : v) x = str();
    return v;
  }
  template <typename T>
  std::vector<T> pod() {
    const uint64_t n = u64();
    if (n > s_.size() / sizeof(T) + 1) bad();
    need(n * sizeof(T));
    std::vector<T> v(n);
    std::memcpy(v.data(), s_.data() + p_, n * sizeof(T));
    p_ += n * sizeof(T);
    return v;
  }
  size_t count(uint64_t limit) {
    const uint64_t n = u64();
    if (n > limit) bad();
    return (size_t)n;
  }

 private:
  void need(uint64_t n) {
    if (p_ + n > s_.size() || p_ + n < p_) bad();
  }
  [[noreturn]] void bad() { throw GlsError(GLS_E_FORMAT, "inter.dat: damaged compiled section"); }
  const std::string& s_;
  size_t p_;
};

}  // namespace

std::string inter_file_text(const Library& lib, const Netlist& nl, const Program& pg,
                            const DesignTexts& texts) {
  std::string out = inter_archive_text(lib, nl, pg);
  BinOut b;
  // library
  b.u64(lib.types.size());
  for (auto& t : lib.types) {
    b.str(t.name);
    b.strs(t.ports);
    b.strs(t.inputs);
    b.strs(t.outputs);
    b.u64(t.prims.size());
    for (auto& p : t.prims) {
      b.i64(p.op);
      b.str(p.out);
      b.strs(p.ins);
    }
    b.strs(t.supply0);
    b.strs(t.supply1);
  }
  b.strs(lib.ignored);
  // netlist
  b.str(nl.top);
  b.strs(nl.net_names);
  b.pod(nl.net_is_input);
  b.pod(nl.const_val);
  b.pod(nl.const_hidden);
  b.pod(nl.alias_of);
  b.u64(nl.insts.size());
  for (auto& in : nl.insts) {
    b.str(in.name);
    b.i64(in.type);
    b.u64(in.in_pins.size());
    for (auto& p : in.in_pins) {
      b.str(p.first);
      b.i64(p.second);
    }
    b.u64(in.out_pins.size());
    for (auto& p : in.out_pins) {
      b.str(p.first);
      b.i64(p.second);
    }
    b.strs(in.unknown_out);
    b.u64(in.delays.size());
    for (auto& d : in.delays) {
      b.i64(d.edge);
      b.i64(d.in_net);
      b.i64(d.out_net);
      b.f64(d.rise);
      b.f64(d.fall);
    }
  }
  // program compiled against the declared inputs
  b.strs(pg.net_names);
  b.pod(pg.role);
  b.pod(pg.const_val);
  b.pod(pg.in_saif);
  b.u64(pg.gates.size());
  for (auto& g : pg.gates) {
    b.i64(g.inst);
    b.i64(g.out_net);
    b.i64(g.k);
    for (int j = 0; j < kMaxK; ++j) b.i64(g.in_net[j]);
    b.i64(g.lut);
    for (int j = 0; j < kMaxK; ++j) {
      b.u64(g.recs[j].size());
      for (auto& r : g.recs[j]) {
        b.i64(r.kind);
        b.f64(r.rise);
        b.f64(r.fall);
      }
    }
  }
  b.pod(pg.level_start);
  b.u64(pg.luts.size());
  for (auto& t : pg.luts) b.pod(t);
  b.pod(pg.lut_k);
  b.i64(pg.n_netlist_nets);
  b.u64(pg.n_pins);
  b.str(texts.vlib);
  b.str(texts.gv);
  b.str(texts.sdf);
  const size_t at = out.size();
  out += kMagic;
  char len[48];
  std::snprintf(len, sizeof len, "%llu\n", (unsigned long long)b.text().size());
  out += len;
  out += b.text();
  std::snprintf(len, sizeof len, "\nGLS-CSR-AT %020llu\n", (unsigned long long)at);
  out += len;
  return out;
}

bool read_inter_section(const std::string& text, Library& lib, Netlist& nl, Program& pg,
                        DesignTexts& texts) {
  if (text.size() < kTrailer || text.compare(text.size() - kTrailer, 12, "\nGLS-CSR-AT ") != 0)
    return false;
  const uint64_t at = std::strtoull(text.c_str() + text.size() - kTrailer + 12, nullptr, 10);
  const size_t body_end = text.size() - kTrailer;
  if (at >= body_end || text.compare(at, sizeof(kMagic) - 1, kMagic) != 0)
    throw GlsError(GLS_E_FORMAT, "inter.dat: damaged compiled section");
  size_t p = at + sizeof(kMagic) - 1;
  const size_t eol = text.find('\n', p);
  if (eol == std::string::npos || eol >= body_end)
    throw GlsError(GLS_E_FORMAT, "inter.dat: damaged compiled section");
  const uint64_t n = std::strtoull(text.substr(p, eol - p).c_str(), nullptr, 10);
  p = eol + 1;
  if (p + n != body_end) throw GlsError(GLS_E_FORMAT, "inter.dat: truncated compiled section");
  BinIn b(text, p);
  const size_t lim = text.size();
  lib = Library();
  const size_t nt = b.count(lim);
  for (size_t i = 0; i < nt; ++i) {
    CellType t;
    t.name = b.str();
    t.ports = b.strs();
    t.inputs = b.strs();
    t.outputs = b.strs();
    const size_t np = b.count(lim);
    for (size_t q = 0; q < np; ++q) {
      Prim pr;
      pr.op = (int)b.i64();
      if (pr.op < 0 || pr.op > P_MUX2) throw GlsError(GLS_E_FORMAT, "inter.dat: bad primitive");
      pr.out = b.str();
      pr.ins = b.strs();
      t.prims.push_back(std::move(pr));
    }
    t.supply0 = b.strs();
    t.supply1 = b.strs();
    lib.add(std::move(t));
  }
  lib.ignored = b.strs();
  nl = Netlist();
  nl.top = b.str();
  nl.net_names = b.strs();
  nl.net_is_input = b.pod<uint8_t>();
  nl.const_val = b.pod<int>();
  nl.const_hidden = b.pod<uint8_t>();
  nl.alias_of = b.pod<int>();
  const size_t N = nl.net_names.size();
  if (nl.net_is_input.size() != N || nl.const_val.size() != N || nl.const_hidden.size() != N ||
      nl.alias_of.size() != N)
    throw GlsError(GLS_E_FORMAT, "inter.dat: net tables disagree");
  for (size_t i = 0; i < N; ++i) nl.net_ids.emplace(nl.net_names[i], (int)i);
  auto net_ok = [&](int64_t x) {
    if (x < 0 || (size_t)x >= N) throw GlsError(GLS_E_FORMAT, "inter.dat: net out of range");
    return (int)x;
  };
  const size_t ni = b.count(lim);
  nl.insts.resize(ni);
  for (auto& in : nl.insts) {
    in.name = b.str();
    in.type = (int)b.i64();
    if (in.type < 0 || (size_t)in.type >= lib.types.size())
      throw GlsError(GLS_E_FORMAT, "inter.dat: cell type out of range");
    const size_t a = b.count(lim);
    for (size_t q = 0; q < a; ++q) {
      std::string pin = b.str();
      in.in_pins.emplace_back(std::move(pin), net_ok(b.i64()));
    }
    const size_t o = b.count(lim);
    for (size_t q = 0; q < o; ++q) {
      std::string pin = b.str();
      in.out_pins.emplace_back(std::move(pin), net_ok(b.i64()));
    }
    in.unknown_out = b.strs();
    const size_t nd = b.count(lim);
    for (size_t q = 0; q < nd; ++q) {
      IoPath d;
      d.edge = (int)b.i64();
      d.in_net = (int)b.i64();
      d.out_net = (int)b.i64();
      d.rise = b.f64();
      d.fall = b.f64();
      in.delays.push_back(d);
    }
  }
  pg = Program();
  pg.net_names = b.strs();
  pg.role = b.pod<uint8_t>();
  pg.const_val = b.pod<int>();
  pg.in_saif = b.pod<uint8_t>();
  const size_t PN = pg.net_names.size();
  if (pg.role.size() != PN || pg.const_val.size() != PN || pg.in_saif.size() != PN)
    throw GlsError(GLS_E_FORMAT, "inter.dat: program net tables disagree");
  const size_t ng = b.count(lim);
  pg.gates.resize(ng);
  for (auto& g : pg.gates) {
    g.inst = (int)b.i64();
    g.out_net = (int)b.i64();
    g.k = (int)b.i64();
    if (g.k < 1 || g.k > kMaxK || g.out_net < 0 || (size_t)g.out_net >= PN || g.inst < 0 ||
        (size_t)g.inst >= ni)
      throw GlsError(GLS_E_FORMAT, "inter.dat: bad gate");
    for (int j = 0; j < kMaxK; ++j) g.in_net[j] = (int)b.i64();
    for (int j = 0; j < g.k; ++j)
      if (g.in_net[j] < 0 || (size_t)g.in_net[j] >= PN) throw GlsError(GLS_E_FORMAT, "inter.dat: bad gate input");
    g.lut = (int)b.i64();
    for (int j = 0; j < kMaxK; ++j) {
      const size_t nr = b.count(lim);
      for (size_t q = 0; q < nr; ++q) {
        CompiledGate::Rec r;
        r.kind = (int)b.i64();
        r.rise = b.f64();
        r.fall = b.f64();
        g.recs[j].push_back(r);
      }
    }
  }
  pg.level_start = b.pod<uint32_t>();
  const size_t nlut = b.count(lim);
  for (size_t q = 0; q < nlut; ++q) pg.luts.push_back(b.pod<uint8_t>());
  pg.lut_k = b.pod<int>();
  pg.n_netlist_nets = (int)b.i64();
  pg.n_pins = b.u64();
  texts.vlib = b.str();
  texts.gv = b.str();
  texts.sdf = b.str();
  if (pg.level_start.empty() || pg.level_start.back() != ng || pg.lut_k.size() != pg.luts.size())
    throw GlsError(GLS_E_FORMAT, "inter.dat: bad level table");
  for (auto& g : pg.gates)
    if (g.lut < 0 || (size_t)g.lut >= pg.luts.size() || pg.luts[g.lut].size() != (size_t)(1u << (2 * g.k)))
      throw GlsError(GLS_E_FORMAT, "inter.dat: bad truth table reference");
  return true;
}

}  // namespace gls
