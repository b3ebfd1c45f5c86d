#include "formats.h"

#include <algorithm>
#include <array>
#include <exception>
#include <thread>
#include <cctype>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <unordered_map>

#include "../../include/gls_status.h"

namespace gls {

std::string read_file(const std::string& path) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) throw GlsError(GLS_E_IO, "cannot open " + path);
  std::string s;
  std::fseek(f, 0, SEEK_END);
  long n = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  if (n > 0) {
    s.resize((size_t)n);
    size_t got = std::fread(&s[0], 1, (size_t)n, f);
    s.resize(got);
  }
  std::fclose(f);
  return s;
}

void write_file(const std::string& path, const std::string& data) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) throw GlsError(GLS_E_IO, "cannot write " + path);
  size_t w = std::fwrite(data.data(), 1, data.size(), f);
  std::fclose(f);
  if (w != data.size()) throw GlsError(GLS_E_IO, "short write to " + path);
}

// strtok-style split: every char of `delims` separates, empty tokens dropped
// (the reference's PreProcess::Split, src/PreProcess.cpp:1466-1484).
static std::vector<std::string> split(const std::string& s, const char* delims) {
  std::vector<std::string> r;
  size_t i = 0, n = s.size();
  while (i < n) {
    while (i < n && std::strchr(delims, s[i]) && s[i] != 0) ++i;
    size_t j = i;
    while (j < n && !(std::strchr(delims, s[j]) && s[j] != 0)) ++j;
    if (j > i) r.push_back(s.substr(i, j - i));
    i = j;
  }
  return r;
}

static std::string lower(std::string s) {
  for (auto& c : s) c = (char)std::tolower((unsigned char)c);
  return s;
}
static std::string upper(std::string s) {
  for (auto& c : s) c = (char)std::toupper((unsigned char)c);
  return s;
}
static std::string trim_spaces(const std::string& s) {
  size_t a = s.find_first_not_of(' ');
  if (a == std::string::npos) return "";
  size_t b = s.find_last_not_of(' ');
  return s.substr(a, b - a + 1);
}

static std::vector<std::string> lines_of(const std::string& text) {
  std::vector<std::string> out;
  size_t p = 0;
  while (p <= text.size()) {
    size_t q = text.find('\n', p);
    if (q == std::string::npos) {
      if (p < text.size()) out.push_back(text.substr(p));
      break;
    }
    out.push_back(text.substr(p, q - p));
    p = q + 1;
  }
  return out;
}

// ---------------------------------------------------------------- vlib ----
// src/PreProcess.cpp:42-648.
Library parse_vlib_text(const std::string& text) {
  Library lib;
  auto lines = lines_of(text);
  const char* D = " \t\r\n,;()";
  int const_counter = 0;
  for (size_t li = 0; li < lines.size(); ++li) {
    auto tmp = split(lines[li], D);
    if (tmp.empty()) continue;
    if (lower(tmp[0]) != "module") continue;
    if (tmp.size() < 2) throw GlsError(GLS_E_FORMAT, "vlib: module without a name");
    CellType ct;
    ct.name = tmp[1];
    bool ignored = false, closed = false;
    for (++li; li < lines.size(); ++li) {
      auto t = split(lines[li], D);
      if (t.empty()) continue;
      std::string kw = lower(t[0]);
      if (kw == "input" || kw == "output" || kw == "wire" || kw == "reg" || kw == "inout") {
        if (t.size() < 2) throw GlsError(GLS_E_FORMAT, "vlib: empty declaration in " + ct.name);
        std::vector<std::string> names;
        if (t[1].find('[') != std::string::npos) {
          if (t.size() < 3) throw GlsError(GLS_E_FORMAT, "vlib: bus declaration without a name");
          names.push_back(t[2]);
        } else {
          names.assign(t.begin() + 1, t.end());
        }
        for (auto& n : names) {
          if (kw == "input" || kw == "output" || kw == "inout") ct.ports.push_back(n);
          if (kw == "input") ct.inputs.push_back(n);
          if (kw == "output") ct.outputs.push_back(n);
        }
      } else if (kw == "and" || kw == "or" || kw == "xor" || kw == "xnor" || kw == "nor" ||
                 kw == "nand" || kw == "buf" || kw == "not" || kw == "udp_xbuf" ||
                 kw == "udp_mux2") {
        Prim p;
        static const char* ops[] = {"and", "or", "xor", "xnor", "nor", "nand",
                                    "buf", "not", "udp_xbuf", "udp_mux2"};
        for (int o = 0; o < 10; ++o)
          if (kw == ops[o]) p.op = o;
        size_t first = (p.op == P_XBUF || p.op == P_MUX2) ? 2 : 1;  // UDPs carry an instance name
        if (t.size() < first + 2) throw GlsError(GLS_E_FORMAT, "vlib: primitive without pins in " + ct.name);
        if (p.op == P_XBUF && t.size() != 5) throw GlsError(GLS_E_FORMAT, "vlib: udp_xbuf needs 2 inputs");
        if (p.op == P_MUX2 && t.size() != 6) throw GlsError(GLS_E_FORMAT, "vlib: udp_mux2 needs 3 inputs");
        p.out = t[first];
        for (size_t i = first + 1; i < t.size(); ++i) {
          if (t[i] == "1'b1" || t[i] == "1'b0") {
            std::string nm = "zc_const" + std::to_string(const_counter++);
            (t[i] == "1'b1" ? ct.supply1 : ct.supply0).push_back(nm);
            p.ins.push_back(nm);
          } else {
            p.ins.push_back(t[i]);
          }
        }
        ct.prims.push_back(p);
      } else if (kw == "supply0") {
        for (size_t i = 1; i < t.size(); ++i) ct.supply0.push_back(t[i]);
      } else if (kw == "supply1") {
        for (size_t i = 1; i < t.size(); ++i) ct.supply1.push_back(t[i]);
      } else if (kw == "udp_dff" || kw == "udp_tlat" || kw == "always") {
        ignored = true;
        break;
      } else if (kw == "endmodule") {
        closed = true;
        break;
      }
    }
    if (ignored) {
      lib.ignored.push_back(ct.name);
      continue;
    }
    if (!closed) throw GlsError(GLS_E_FORMAT, "vlib: module " + ct.name + " has no endmodule");
    lib.add(std::move(ct));
  }
  return lib;
}

Library parse_vlib_file(const std::string& path) { return parse_vlib_text(read_file(path)); }

// ------------------------------------------------------------------ gv ----
// src/PreProcess.cpp:650-1274, restricted to flat single-module netlists
// with named .pin(net) connections, scalar and bus nets, {..}
// concatenations, 1'b0/1'b1 ties and `assign`.
static std::string strip_tail(const std::string& s) {
  if (!s.empty() && (s.back() == ',' || s.back() == ';')) return s.substr(0, s.size() - 1);
  return s;
}

static void parse_range(const std::string& tok, int& l, int& r) {
  size_t a = tok.find('['), c = tok.find(':'), b = tok.find(']');
  if (a == std::string::npos || c == std::string::npos || b == std::string::npos)
    throw GlsError(GLS_E_FORMAT, "gv: bad range " + tok);
  l = std::stoi(tok.substr(a + 1, c - a - 1));
  r = std::stoi(tok.substr(c + 1, b - c - 1));
}

// The text is read in three steps: a serial scan cuts it into statements
// (cheap: it looks only at each line's first word and last character); the
// statements are then tokenised in parallel into net names and pins; a
// serial pass finally numbers the nets in text order and applies every
// check that depends on earlier statements (undeclared nets).  Net numbers,
// instance order and the first error reported are those of a one-thread
// read.
namespace {

struct GvStmt {
  uint8_t kind = 0;                // 0 nothing, 1 declaration, 2 assign, 3 instance
  bool input = false;              // declaration: `input`
  std::vector<std::string> names;  // declared nets; assign: left, right; instance: net per pin bit
  std::vector<std::string> pin_bits;
  std::vector<uint8_t> pin_in;     // instance: 1 input pin, 0 output pin
  Instance inst;                   // instance: name, type, pins, unconnected outputs
  size_t n_ok = 0;                 // instance: pin bits read before `err`
  // instance pin bits whose net was not numbered before this round (new
  // nets, constants): (pin bit, 1 input / 0 output, index in in_pins/out_pins)
  std::vector<std::array<uint32_t, 3>> todo;
  std::exception_ptr err;
};

inline bool gv_delim(char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\n'; }

// First word of a line, lower-cased, compared with a keyword.
bool first_word_is(const std::string& text, size_t a, size_t b, const char* kw) {
  while (a < b && gv_delim(text[a])) ++a;
  size_t e = a;
  while (e < b && !gv_delim(text[e])) ++e;
  const size_t n = std::strlen(kw);
  if (e - a != n) return false;
  for (size_t k = 0; k < n; ++k)
    if ((char)std::tolower((unsigned char)text[a + k]) != kw[k]) return false;
  return true;
}

// Pins of an instance, numbered where the net already has a number (nets
// are never renumbered, so a number found now is the one a one-thread read
// finds); the others are left to the serial pass.
void place_pins(const Netlist& nl, GvStmt& R) {
  for (size_t b = 0; b < R.n_ok; ++b) {
    const std::string& nn = R.names[b];
    const int id = (nn == "1'b1" || nn == "1'b0") ? -1 : nl.net(nn);
    auto& pins = R.pin_in[b] ? R.inst.in_pins : R.inst.out_pins;
    if (id < 0) R.todo.push_back({(uint32_t)b, (uint32_t)R.pin_in[b], (uint32_t)pins.size()});
    pins.push_back({std::move(R.pin_bits[b]), id});
  }
}

void parse_gv_stmt(const std::string& stmt, const Library& lib, GvStmt& R) {
  auto t = split(stmt, " \t\r\n");
  if (t.empty()) return;
  std::string k0 = lower(t[0]);
  if (k0 == "input" || k0 == "output" || k0 == "wire" || k0 == "inout") {
    R.kind = 1;
    R.input = k0 == "input";
    if (t.size() < 2) throw GlsError(GLS_E_FORMAT, "gv: empty declaration");
    if (t[1].find('[') != std::string::npos) {
      if (t.size() < 3) throw GlsError(GLS_E_FORMAT, "gv: bus declaration without a name");
      int l, r;
      parse_range(t[1], l, r);
      // a bus declaration may list several names: input [3:0] a, b;
      for (size_t ni = 2; ni < t.size(); ++ni) {
        std::string nm = strip_tail(t[ni]);
        if (nm.empty()) continue;
        int step = l > r ? -1 : 1;
        for (int i = l;; i += step) {
          R.names.push_back(nm + "[" + std::to_string(i) + "]");
          if (i == r) break;
        }
      }
    } else {
      for (size_t i = 1; i < t.size(); ++i) {
        std::string nm = strip_tail(t[i]);
        if (!nm.empty()) R.names.push_back(std::move(nm));
      }
    }
    return;
  }
  if (k0 == "assign") {
    // assign left = right;  (any spacing)
    R.kind = 2;
    std::string body;
    for (size_t i = 1; i < t.size(); ++i) body += t[i];
    if (!body.empty() && body.back() == ';') body.pop_back();
    size_t eq = body.find('=');
    if (eq == std::string::npos) throw GlsError(GLS_E_FORMAT, "gv: assign without '='");
    R.names.push_back(body.substr(0, eq));
    R.names.push_back(body.substr(eq + 1));
    return;
  }
  // instance statement
  int ty = lib.find(t[0]);
  if (ty < 0) return;  // the reference skips unknown modules with a warning
  R.kind = 3;
  const CellType& ct = lib.types[ty];
  std::string body;
  for (size_t i = 1; i < t.size(); ++i) body += t[i];
  size_t lp = body.find('(');
  if (lp == std::string::npos) throw GlsError(GLS_E_FORMAT, "gv: instance without connections");
  Instance& inst = R.inst;
  inst.name = body.substr(0, lp);
  inst.type = ty;
  std::vector<std::string> used_ports;
  // connections: .pin(net) or .pin({a,b[3:0],1'b0})
  size_t p = lp + 1;
  while ((p = body.find('.', p)) != std::string::npos) {
    size_t po = body.find('(', p);
    if (po == std::string::npos) throw GlsError(GLS_E_FORMAT, "gv: bad connection in " + inst.name);
    std::string pin = body.substr(p + 1, po - p - 1);
    int depth = 1;
    size_t q = po + 1;
    while (q < body.size() && depth > 0) {
      if (body[q] == '(') depth++;
      else if (body[q] == ')') depth--;
      if (depth > 0) ++q;
    }
    std::string conn = body.substr(po + 1, q - po - 1);
    p = q;
    auto it = std::find(ct.ports.begin(), ct.ports.end(), pin);
    if (it == ct.ports.end()) throw GlsError(GLS_E_FORMAT, "gv: " + ct.name + " has no pin " + pin);
    bool is_in = std::find(ct.inputs.begin(), ct.inputs.end(), pin) != ct.inputs.end();
    bool is_out = std::find(ct.outputs.begin(), ct.outputs.end(), pin) != ct.outputs.end();
    if (!is_in && !is_out) throw GlsError(GLS_E_FORMAT, "gv: pin " + pin + " must be input or output");
    used_ports.push_back(pin);
    std::vector<std::string> nets;
    if (!conn.empty() && conn[0] == '{') {
      std::string inner = conn.substr(1, conn.size() - 2);
      for (auto& part : split(inner, ",")) {
        if (part.find('[') != std::string::npos && part.find(':') != std::string::npos) {
          int l, r;
          parse_range(part, l, r);
          std::string base = part.substr(0, part.find('['));
          int step = l > r ? -1 : 1;
          for (int i = l;; i += step) {
            nets.push_back(base + "[" + std::to_string(i) + "]");
            if (i == r) break;
          }
        } else {
          nets.push_back(part);
        }
      }
    } else {
      nets.push_back(conn);
    }
    for (size_t bi = 0; bi < nets.size(); ++bi) {
      R.pin_bits.push_back(nets.size() > 1 ? pin + "[" + std::to_string(bi) + "]" : pin);
      R.names.push_back(std::move(nets[bi]));
      R.pin_in.push_back(is_in ? 1 : 0);
      R.n_ok = R.names.size();
    }
  }
  for (auto& op : ct.outputs)
    if (std::find(used_ports.begin(), used_ports.end(), op) == used_ports.end())
      inst.unknown_out.push_back(op);
}

}  // namespace

Netlist parse_gv_text(const std::string& text, const Library& lib, int threads) {
  Netlist nl;
  // 1. statements: spans of trimmed lines; a statement is " " + line for
  // each of its lines, ending with the line that ends in ';'
  std::vector<std::pair<size_t, size_t>> spans;  // trimmed lines of all statements
  std::vector<size_t> first;                     // statement k: spans[first[k], first[k+1])
  {
    bool in_module = false, header = false;
    size_t open = 0;  // spans of the statement being assembled
    const size_t n = text.size();
    size_t p = 0;
    while (p < n) {
      size_t q = text.find('\n', p);
      if (q == std::string::npos) q = n;
      size_t a = p, b = q;
      p = q + 1;
      while (a < b && text[a] == ' ') ++a;
      while (b > a && text[b - 1] == ' ') --b;
      if (a == b) continue;
      size_t w = a;
      while (w < b && gv_delim(text[w])) ++w;
      if (w == b) continue;  // no word
      const bool semi = text[b - 1] == ';';
      if (first_word_is(text, a, b, "module")) {
        in_module = true;
        auto tmp = split(text.substr(a, b - a), " \t\r\n");
        if (tmp.size() > 1) {
          const std::string& nm = tmp[1];
          size_t k = nm.find('(');
          nl.top = k == std::string::npos ? nm : nm.substr(0, k);
        }
        // the header's port list may span lines; it declares nothing
        if (!semi) {
          header = true;
          spans.resize(spans.size() - open);
          open = 0;
        }
        continue;
      }
      if (first_word_is(text, a, b, "endmodule")) break;
      if (!in_module) continue;
      if (header) {
        if (semi) header = false;
        continue;
      }
      spans.push_back({a, b - a});
      ++open;
      if (!semi) continue;
      first.push_back(spans.size() - open);
      open = 0;
    }
    spans.resize(spans.size() - open);  // an unterminated statement is dropped
    first.push_back(spans.size());
  }
  const size_t n_stmt = first.size() - 1;
  // 2 + 3, in rounds that bound the memory held by tokenised statements
  threads = host_threads(threads);
  if (text.size() < parallel_min_bytes()) threads = 1;
  int hidden_const = 0;
  auto const_net = [&](const std::string& v) {
    std::string nm = "gv_net_const" + std::to_string(hidden_const++);
    int id = nl.add_net(nm);
    nl.const_val[id] = (v == "1'b1") ? V1 : V0;
    nl.const_hidden[id] = 1;
    return id;
  };
  const size_t round = 1u << 17;
  nl.net_ids.reserve(text.size() / 64 + 16);  // about one net per 64 bytes of text
  nl.insts.reserve(n_stmt);
  std::vector<GvStmt> R;
  for (size_t r0 = 0; r0 < n_stmt; r0 += round) {
    const size_t nr = std::min(round, n_stmt - r0);
    R.clear();
    R.resize(nr);
    auto work = [&](size_t lo, size_t hi) {
      std::string stmt;
      for (size_t k = lo; k < hi; ++k) {
        try {
          stmt.clear();
          for (size_t x = first[r0 + k]; x < first[r0 + k + 1]; ++x) {
            stmt += ' ';
            stmt.append(text, spans[x].first, spans[x].second);
          }
          try {
            parse_gv_stmt(stmt, lib, R[k]);
          } catch (...) {
            R[k].err = std::current_exception();
          }
          if (R[k].kind == 3) place_pins(nl, R[k]);
        } catch (...) {  // out of memory: reported as soon as the serial pass reaches it
          R[k].kind = 0;
          R[k].err = std::current_exception();
        }
      }
    };
    const size_t nt = std::min<size_t>((size_t)threads, (nr + 1023) / 1024);
    if (nt <= 1) {
      work(0, nr);
    } else {
      std::vector<std::thread> th;
      for (size_t w = 0; w < nt; ++w) th.emplace_back(work, nr * w / nt, nr * (w + 1) / nt);
      for (auto& t : th) t.join();
    }
    for (size_t k = 0; k < nr; ++k) {
      GvStmt& S = R[k];
      if (S.kind == 1 && !S.err) {
        for (auto& nm : S.names) {
          int id = nl.add_net(nm);
          if (S.input) nl.net_is_input[id] = 1;
        }
      } else if (S.kind == 2 && !S.err) {
        int lid = nl.add_net(S.names[0]);
        const std::string& right = S.names[1];
        if (right == "1'b1" || right == "1'b0") {
          nl.const_val[lid] = right == "1'b1" ? V1 : V0;
        } else {
          int rid = nl.add_net(right);
          nl.alias_of[lid] = rid;
        }
      } else if (S.kind == 3) {
        for (auto& td : S.todo) {
          const std::string& nn = S.names[td[0]];
          int id;
          if (nn == "1'b1" || nn == "1'b0") {
            id = const_net(nn);
          } else {
            id = nl.net(nn);
            if (id < 0) {
              if (nn.find('[') == std::string::npos)
                throw GlsError(GLS_E_FORMAT, "gv: undeclared net " + nn + " on " + S.inst.name);
              id = nl.add_net(nn);
            }
          }
          (td[1] ? S.inst.in_pins : S.inst.out_pins)[td[2]].second = id;
        }
        if (!S.err) nl.insts.push_back(std::move(S.inst));
      }
      if (S.err) std::rethrow_exception(S.err);
    }
  }
  return nl;
}

Netlist parse_gv_file(const std::string& path, const Library& lib, int threads) {
  return parse_gv_text(read_file(path), lib, threads);
}

// ----------------------------------------------------------------- SDF ----
// src/PreProcess.cpp:1276-1464: TIMESCALE factor (default 1ns = 1e3),
// INSTANCE blocks closed by the first line that trims to ")", typ value of
// each triple times the factor.
//
// A large file is read in parallel stretches.  A stretch starts right after
// a line that trims to ")": such a line ends any open INSTANCE block, so
// every stretch starts outside a block, as the one-thread read would be
// there.  What a stretch cannot know is the TIMESCALE factor in force at its
// start; its delays are kept unscaled up to its own first TIMESCALE line and
// multiplied by the inherited factor when the stretches are joined in order
// (the same single multiplication the one-thread read makes).
namespace {

struct SdfPart {
  std::vector<std::pair<int, IoPath>> paths;  // (instance, path) in text order
  std::vector<uint8_t> scaled;                // path's values already carry a factor
  bool has_factor = false;                    // a TIMESCALE line of this stretch set one
  double factor = 1e3;                        // factor in force at the end, when has_factor
  std::exception_ptr err;
};

// Iterate the lines of text[a, b) as lines_of() cuts them.
template <class F>
void for_lines(const std::string& text, size_t a, size_t b, F&& f) {
  size_t p = a;
  while (p < b) {
    size_t q = text.find('\n', p);
    if (q == std::string::npos || q > b) q = b;
    if (!f(text.substr(p, q - p))) return;
    p = q + 1;
  }
}

void parse_sdf_part(const std::string& text, size_t a, size_t b, const Netlist& nl,
                    const std::unordered_map<std::string, int>& by_name, SdfPart& P) {
  const char* D = " \t\r\n\"'()";
  bool in_block = false;  // inside the INSTANCE block of a known instance
  int cur = -1;
  auto typ = [&](const std::string& tok, uint8_t& sc) -> double {
    auto v = split(tok, ":");
    const std::string& s = v.size() == 3 ? v[1] : (v.empty() ? tok : v[0]);
    char* end = nullptr;
    double x = std::strtod(s.c_str(), &end);
    if (end == s.c_str()) throw GlsError(GLS_E_FORMAT, "sdf: bad delay value " + tok);
    sc = P.has_factor ? 1 : 0;
    return P.has_factor ? x * P.factor : x;
  };
  for_lines(text, a, b, [&](const std::string& raw) {
    std::string line = trim_spaces(raw);
    if (line.empty()) return true;
    if (in_block) {
      if (line == ")") {
        in_block = false;
        return true;
      }
      auto t = split(line, D);
      if (t.empty() || upper(t[0]) != "IOPATH") return true;
      const Instance& inst = nl.insts[cur];
      IoPath d;
      size_t ip, op, vp;
      std::string e = upper(t.size() > 1 ? t[1] : "");
      if (e == "POSEDGE" || e == "NEGEDGE") {
        if (t.size() < 5) throw GlsError(GLS_E_FORMAT, "sdf: short IOPATH in " + inst.name);
        d.edge = e == "POSEDGE" ? 0 : 1;
        ip = 2; op = 3; vp = 4;
      } else {
        if (t.size() < 4) throw GlsError(GLS_E_FORMAT, "sdf: short IOPATH in " + inst.name);
        d.edge = 2;
        ip = 1; op = 2; vp = 3;
      }
      d.in_net = -1;
      for (auto& pp : inst.in_pins)
        if (pp.first == t[ip]) { d.in_net = pp.second; break; }
      d.out_net = -1000000;
      for (auto& pp : inst.out_pins)
        if (pp.first == t[op]) { d.out_net = pp.second; break; }
      if (d.out_net == -1000000) {
        for (size_t k = 0; k < inst.unknown_out.size(); ++k)
          if (inst.unknown_out[k] == t[op]) { d.out_net = -2 - (int)k; break; }
      }
      if (d.out_net == -1000000)
        throw GlsError(GLS_E_FORMAT, "sdf: instance " + inst.name + " has no output pin " + t[op]);
      uint8_t sc = 0;
      d.rise = typ(t[vp], sc);
      d.fall = t.size() > vp + 1 ? typ(t[vp + 1], sc) : d.rise;
      P.paths.push_back({cur, d});
      P.scaled.push_back(sc);
      return true;
    }
    auto tmp = split(line, D);
    if (tmp.empty()) return true;
    std::string k0 = upper(tmp[0]);
    if (k0 == "TIMESCALE" && tmp.size() > 1) {
      const std::string& ts = tmp[1];
      double f = 0;
      if (ts == "1ns") f = 1e3;
      else if (ts == "1ps") f = 1.0;
      else if (ts == "1us") f = 1e6;
      else if (ts == "1ms") f = 1e9;
      else if (ts == "1s") f = 1e12;
      // other spellings keep the current factor, as the reference does
      if (f != 0) {
        P.factor = f;
        P.has_factor = true;
      }
      return true;
    }
    if (k0 != "INSTANCE" || tmp.size() < 2) return true;
    auto it = by_name.find(tmp[1]);
    if (it == by_name.end()) return true;
    cur = it->second;
    in_block = true;
    return true;
  });
}

}  // namespace

void parse_sdf_text(const std::string& text, Netlist& nl, int threads) {
  std::unordered_map<std::string, int> by_name;
  by_name.reserve(nl.insts.size());
  for (size_t i = 0; i < nl.insts.size(); ++i) by_name[nl.insts[i].name] = (int)i;
  const size_t n = text.size();
  threads = host_threads(threads);
  if (n < parallel_min_bytes()) threads = 1;
  // stretch starts: the line after a line that trims to ")"
  std::vector<size_t> cut{0};
  for (int i = 1; i < threads; ++i) {
    // first line start at or after the target
    size_t p = std::max(n / (size_t)threads * (size_t)i, cut.back());
    if (p > 0 && text[p - 1] != '\n') {
      p = text.find('\n', p);
      if (p == std::string::npos) break;
      ++p;
    }
    size_t found = std::string::npos;
    while (p < n) {
      size_t q = text.find('\n', p);
      if (q == std::string::npos) break;
      if (trim_spaces(text.substr(p, q - p)) == ")") {
        found = q + 1;
        break;
      }
      p = q + 1;
    }
    if (found == std::string::npos || found >= n) break;
    if (found > cut.back()) cut.push_back(found);
  }
  cut.push_back(n);
  const size_t np = cut.size() - 1;
  std::vector<SdfPart> parts(np);
  auto run_part = [&](size_t i) {
    try {
      parse_sdf_part(text, cut[i], cut[i + 1], nl, by_name, parts[i]);
    } catch (...) {
      parts[i].err = std::current_exception();
    }
  };
  if (np == 1) {
    run_part(0);
  } else {
    std::vector<std::thread> th;
    for (size_t i = 0; i < np; ++i) th.emplace_back(run_part, i);
    for (auto& t : th) t.join();
  }
  // the first error in text order is the one a one-thread read reports
  for (auto& P : parts)
    if (P.err) std::rethrow_exception(P.err);
  double factor = 1e3;
  for (auto& P : parts) {
    for (size_t k = 0; k < P.paths.size(); ++k) {
      IoPath& d = P.paths[k].second;
      if (!P.scaled[k]) {
        d.rise = d.rise * factor;
        d.fall = d.fall * factor;
      }
      nl.insts[P.paths[k].first].delays.push_back(d);
    }
    if (P.has_factor) factor = P.factor;
  }
}

void parse_sdf_file(const std::string& path, Netlist& nl, int threads) {
  parse_sdf_text(read_file(path), nl, threads);
}

// -------------------------------------------------------------- digest ----
namespace {
struct Fnv {
  uint64_t h = 1469598103934665603ull;
  void bytes(const void* p, size_t n) {
    const unsigned char* b = (const unsigned char*)p;
    for (size_t i = 0; i < n; ++i) h = (h ^ b[i]) * 1099511628211ull;
  }
  void u64(uint64_t x) { bytes(&x, 8); }
  void str(const std::string& s) {
    u64(s.size());
    bytes(s.data(), s.size());
  }
  void f64(double x) {
    uint64_t b;
    std::memcpy(&b, &x, 8);
    u64(b);
  }
};
}  // namespace

uint64_t netlist_digest(const Netlist& nl) {
  Fnv f;
  f.str(nl.top);
  f.u64(nl.net_names.size());
  for (size_t i = 0; i < nl.net_names.size(); ++i) {
    f.str(nl.net_names[i]);
    f.u64(nl.net_is_input[i]);
    f.u64((uint64_t)(int64_t)nl.const_val[i]);
    f.u64(nl.const_hidden[i]);
    f.u64((uint64_t)(int64_t)nl.alias_of[i]);
  }
  f.u64(nl.insts.size());
  for (const Instance& in : nl.insts) {
    f.str(in.name);
    f.u64((uint64_t)(int64_t)in.type);
    for (auto* pins : {&in.in_pins, &in.out_pins}) {
      f.u64(pins->size());
      for (auto& p : *pins) {
        f.str(p.first);
        f.u64((uint64_t)(int64_t)p.second);
      }
    }
    f.u64(in.unknown_out.size());
    for (auto& u : in.unknown_out) f.str(u);
    f.u64(in.delays.size());
    for (const IoPath& d : in.delays) {
      f.u64((uint64_t)(int64_t)d.edge);
      f.u64((uint64_t)(int64_t)d.in_net);
      f.u64((uint64_t)(int64_t)d.out_net);
      f.f64(d.rise);
      f.f64(d.fall);
    }
  }
  return f.h;
}

// ----------------------------------------------------------------- VCD ----
static int vcd_code(char c) {
  switch (c) {
    case '0': return V0;
    case '1': return V1;
    case 'x': case 'X': return VX;
    case 'z': case 'Z': return VZ;
    default: return -1;
  }
}

// Stimulus times are 32-bit (the reference's device path keeps u32 times,
// src/simulate.cu:2479-2482); the simulation moves every time chunk to its
// own time base, so any time below 2^32 is simulated.
static const uint64_t kMaxTime = (1ull << 32) - 1;

namespace {

struct VcdVar {
  std::vector<int> sig;
  int lindex = -1, rindex = -1;
  bool vec = false;
};

// Declarations of a VCD: identifier -> signals, which signals are bus bits,
// and where the value-change body starts.
struct VcdHeader {
  std::unordered_map<std::string, std::vector<VcdVar>> by_id;
  std::vector<uint8_t> is_vec_bit;
  size_t body = 0;
  // identifiers of up to 3 printable characters (the usual case) map to
  // their declarations through a flat table: no allocation or hashing per
  // value change
  std::vector<int> fast;                       // base-95 code -> index into slots, or -1
  std::vector<const std::vector<VcdVar>*> slots;
  static constexpr size_t kFast = 95 * 95 * 95 + 95 * 95 + 95;
  static long code(const char* c, size_t len) {
    if (len == 0 || len > 3) return -1;
    long x = 0;
    for (size_t i = 0; i < len; ++i) {
      const int d = (unsigned char)c[i] - 32;
      if (d < 1 || d > 94) return -1;
      x = x * 95 + d;
    }
    return x;
  }
  const std::vector<VcdVar>* find(const char* c, size_t len) const {
    const long k = code(c, len);
    if (k >= 0) {
      const int i = fast[(size_t)k];
      return i < 0 ? nullptr : slots[(size_t)i];
    }
    auto it = by_id.find(std::string(c, len));
    return it == by_id.end() ? nullptr : &it->second;
  }
};

// Value changes of one stretch of the body, per signal in file order,
// before de-duplication; first/last timestamp seen in the stretch.
struct VcdRaw {
  std::vector<std::vector<uint32_t>> t;
  std::vector<std::vector<uint8_t>> v;
  bool has_stamp = false;
  uint64_t first_stamp = 0, last_stamp = 0;
};

VcdHeader parse_vcd_header(const std::string& s, Stimulus& st) {
  VcdHeader H;
  size_t p = 0, n = s.size();
  auto skip_ws = [&]() { while (p < n && std::isspace((unsigned char)s[p])) ++p; };
  auto word = [&]() {
    skip_ws();
    size_t a = p;
    while (p < n && !std::isspace((unsigned char)s[p])) ++p;
    return s.substr(a, p - a);
  };
  bool scope_seen = false;
  for (;;) {
    std::string w = word();
    if (w.empty()) throw GlsError(GLS_E_FORMAT, "vcd: missing $enddefinitions");
    if (w == "$enddefinitions") {
      while (!(w = word()).empty() && w != "$end") {}
      break;
    }
    if (w == "$timescale") {
      std::string acc;
      while (!(w = word()).empty() && w != "$end") acc += w;
      size_t k = 0;
      while (k < acc.size() && std::isdigit((unsigned char)acc[k])) ++k;
      int mult = k ? std::stoi(acc.substr(0, k)) : 1;
      std::string unit = acc.substr(k);
      long long ps = 1;
      if (unit == "ps") ps = 1;
      else if (unit == "ns") ps = 1000;
      else if (unit == "us") ps = 1000000;
      else if (unit == "fs") throw GlsError(GLS_E_UNSUPPORTED, "vcd: fs timescale unsupported");
      else throw GlsError(GLS_E_FORMAT, "vcd: bad timescale " + acc);
      st.timescale_ps = (int)(mult * ps);
      continue;
    }
    if (w == "$scope") {
      std::string type = word(), name = word();
      if (!scope_seen) { st.scope = name; scope_seen = true; }
      while (!(w = word()).empty() && w != "$end") {}
      continue;
    }
    if (w == "$var") {
      std::vector<std::string> f;
      while (!(w = word()).empty() && w != "$end") f.push_back(w);
      if (f.size() < 4) throw GlsError(GLS_E_FORMAT, "vcd: short $var");
      int size = std::stoi(f[1]);
      std::string id = f[2], ref = f[3];
      VcdVar v;
      std::string range = f.size() > 4 ? f[4] : "";
      if (ref.find('[') != std::string::npos && range.empty()) {
        range = ref.substr(ref.find('['));
        ref = ref.substr(0, ref.find('['));
      }
      if (!range.empty()) {
        std::string inner = range.substr(1, range.size() - 2);
        size_t c = inner.find(':');
        if (c != std::string::npos) {
          v.lindex = std::stoi(inner.substr(0, c));
          v.rindex = std::stoi(inner.substr(c + 1));
          v.vec = true;
        } else {
          // single bit of a bus: its own scalar signal named ref[i]
          ref = ref + "[" + inner + "]";
        }
      } else if (size > 1) {
        v.lindex = size - 1;
        v.rindex = 0;
        v.vec = true;
      }
      if (v.vec) {
        int step = v.lindex > v.rindex ? -1 : 1;
        for (int i = v.lindex;; i += step) {
          v.sig.push_back((int)st.names.size());
          st.names.push_back(ref + "[" + std::to_string(i) + "]");
          st.waves.emplace_back();
          H.is_vec_bit.push_back(1);
          if (i == v.rindex) break;
        }
      } else {
        v.sig.push_back((int)st.names.size());
        st.names.push_back(ref);
        st.waves.emplace_back();
        H.is_vec_bit.push_back(0);
      }
      H.by_id[id].push_back(v);
      continue;
    }
    if (w[0] == '$') {  // $date $version $comment $upscope ...
      if (w != "$upscope" && w != "$end")
        while (!(w = word()).empty() && w != "$end") {}
      continue;
    }
    throw GlsError(GLS_E_FORMAT, "vcd: unexpected token in header: " + w);
  }
  H.body = p;
  H.fast.assign(VcdHeader::kFast, -1);
  for (auto& kv : H.by_id) {
    const long k = VcdHeader::code(kv.first.data(), kv.first.size());
    if (k < 0) continue;
    H.fast[(size_t)k] = (int)H.slots.size();
    H.slots.push_back(&kv.second);
  }
  return H;
}

// Body tokens of s[a, b): timestamps and value changes, recorded per signal
// in file order (no de-duplication yet).  The stretch starts at time 0 or at
// a timestamp of its own.
void parse_vcd_body(const std::string& s, size_t a, size_t b, const VcdHeader& H, VcdRaw& R) {
  const size_t nsig = H.is_vec_bit.size();
  R.t.assign(nsig, {});
  R.v.assign(nsig, {});
  size_t p = a;
  const size_t n = b;
  auto skip_ws = [&]() { while (p < n && std::isspace((unsigned char)s[p])) ++p; };
  auto word = [&]() {
    skip_ws();
    size_t q = p;
    while (p < n && !std::isspace((unsigned char)s[p])) ++p;
    return s.substr(q, p - q);
  };
  uint64_t cur_t = 0;
  auto put = [&](int sig, uint8_t v) {
    R.t[sig].push_back((uint32_t)cur_t);
    R.v[sig].push_back(v);
  };
  while (p < n) {
    skip_ws();
    if (p >= n) break;
    char c = s[p];
    if (c == '#') {
      ++p;
      uint64_t t = 0;
      size_t q = p;
      while (p < n && std::isdigit((unsigned char)s[p])) {
        t = t * 10 + (uint64_t)(s[p++] - '0');
        if (t > kMaxTime) throw GlsError(GLS_E_RANGE, "vcd: time of 2^32 or more");
      }
      if (p == q) throw GlsError(GLS_E_FORMAT, "vcd: bad timestamp");
      if (t < cur_t) throw GlsError(GLS_E_FORMAT, "vcd: timestamps go backwards");
      cur_t = t;
      if (!R.has_stamp) R.first_stamp = t;
      R.has_stamp = true;
      R.last_stamp = t;
      continue;
    }
    if (c == '$') {
      std::string w = word();  // $dumpvars / $dumpall / $dumpon / $dumpoff / $end / $comment
      if (w == "$comment")
        while (!(w = word()).empty() && w != "$end") {}
      continue;
    }
    if (c == 'b' || c == 'B') {
      ++p;
      size_t q = p;
      while (p < n && !std::isspace((unsigned char)s[p])) ++p;
      std::string bits = s.substr(q, p - q);
      skip_ws();
      const size_t iq = p;
      while (p < n && !std::isspace((unsigned char)s[p])) ++p;
      const std::vector<VcdVar>* vars = H.find(s.data() + iq, p - iq);
      if (!vars) throw GlsError(GLS_E_FORMAT, "vcd: unknown identifier " + s.substr(iq, p - iq));
      for (auto& v : *vars) {
        size_t width = v.sig.size();
        std::string bb = bits;
        if (bb.size() < width) {
          char pad = (bb[0] == 'x' || bb[0] == 'X' || bb[0] == 'z' || bb[0] == 'Z') ? bb[0] : '0';
          bb = std::string(width - bb.size(), pad) + bb;
        }
        for (size_t k = 0; k < width; ++k) {
          int code = vcd_code(bb[bb.size() - width + k]);
          if (code < 0) throw GlsError(GLS_E_FORMAT, "vcd: bad vector value");
          put(v.sig[k], (uint8_t)code);
        }
      }
      continue;
    }
    if (c == 'r' || c == 'R') {  // real values are ignored
      word();
      word();
      continue;
    }
    int code = vcd_code(c);
    if (code < 0) throw GlsError(GLS_E_FORMAT, std::string("vcd: unexpected character '") + c + "'");
    ++p;
    size_t q = p;
    while (p < n && !std::isspace((unsigned char)s[p])) ++p;
    const std::vector<VcdVar>* vars = H.find(s.data() + q, p - q);
    if (!vars) throw GlsError(GLS_E_FORMAT, "vcd: unknown identifier " + s.substr(q, p - q));
    for (auto& v : *vars) {
      if (v.vec) {
        for (int sig : v.sig) put(sig, (uint8_t)code);
      } else {
        put(v.sig[0], (uint8_t)code);
      }
    }
  }
}

// One signal's waveform from its value changes in file order: bus bits are
// de-duplicated starting from Z; several changes at one timestamp, the last
// wins.
void finish_vcd_wave(const std::vector<VcdRaw>& parts, int sig, bool vec_bit, Wave& wv) {
  uint8_t last = VZ;
  size_t total = 0;
  for (const VcdRaw& R : parts) total += R.t[sig].size();
  wv.t.reserve(total);
  wv.v.reserve(total);
  for (const VcdRaw& R : parts) {
    const std::vector<uint32_t>& T = R.t[sig];
    const std::vector<uint8_t>& V = R.v[sig];
    for (size_t i = 0; i < T.size(); ++i) {
      const uint8_t v = V[i];
      if (vec_bit) {
        if (v == last) continue;
        last = v;
      }
      if (!wv.t.empty() && wv.t.back() == T[i]) {
        wv.v.back() = v;
        continue;
      }
      wv.push(T[i], v);
    }
  }
}

}  // namespace

// VCD reader (VCDParser::init/parse/process, src/VCDParser.cpp:244-717).
// The header is read first; the body is cut at line-initial timestamps
// ("\n#", as the reference's threaded parser cuts it, src/VCDParser.cpp:668-683)
// into one stretch per thread, the stretches are tokenised in parallel, and
// each signal's waveform is then assembled from the stretches in order, in
// parallel over signals.  The result equals a single-threaded read: a body
// with a $comment (which could hold "\n#") or any error in a stretch is
// read again by one thread, which then gives the canonical error.
Stimulus parse_vcd_text(const std::string& s, int threads) {
  Stimulus st;
  VcdHeader H = parse_vcd_header(s, st);
  const size_t a = H.body, n = s.size();
  threads = host_threads(threads);
  if (n - a < parallel_min_bytes() || s.find("$comment", a) != std::string::npos) threads = 1;
  std::vector<size_t> cut{a};
  for (int i = 1; i < threads; ++i) {
    size_t target = a + (n - a) / (size_t)threads * (size_t)i;
    size_t k = s.find("\n#", std::max(target, cut.back()));
    if (k == std::string::npos) break;
    if (k + 1 > cut.back()) cut.push_back(k + 1);
  }
  cut.push_back(n);
  const size_t np = cut.size() - 1;
  std::vector<VcdRaw> parts(np);
  std::vector<std::exception_ptr> err(np);
  auto run_part = [&](size_t i) {
    try {
      parse_vcd_body(s, cut[i], cut[i + 1], H, parts[i]);
    } catch (...) {
      err[i] = std::current_exception();
    }
  };
  if (np == 1) {
    parse_vcd_body(s, cut[0], cut[1], H, parts[0]);
  } else {
    std::vector<std::thread> th;
    for (size_t i = 0; i < np; ++i) th.emplace_back(run_part, i);
    for (auto& t : th) t.join();
    bool bad = false;
    uint64_t last = 0;
    for (size_t i = 0; i < np && !bad; ++i) {
      if (err[i]) bad = true;
      else if (parts[i].has_stamp) {
        if (parts[i].first_stamp < last) bad = true;  // timestamps go backwards across a cut
        last = parts[i].last_stamp;
      }
    }
    if (bad) return parse_vcd_text(s, 1);
  }
  const size_t nsig = H.is_vec_bit.size();
  const size_t nthr = std::min<size_t>(np, nsig ? nsig : 1);
  std::vector<std::exception_ptr> ferr(std::max<size_t>(nthr, 1));
  auto finish = [&](size_t w) {
    try {
      for (size_t g = w; g < nsig; g += nthr) finish_vcd_wave(parts, (int)g, H.is_vec_bit[g] != 0, st.waves[g]);
    } catch (...) {
      ferr[w] = std::current_exception();
    }
  };
  if (nthr <= 1) {
    finish(0);
  } else {
    std::vector<std::thread> th;
    for (size_t w = 0; w < nthr; ++w) th.emplace_back(finish, w);
    for (auto& t : th) t.join();
  }
  for (auto& e : ferr)
    if (e) std::rethrow_exception(e);
  return st;
}

Stimulus parse_vcd_text(const std::string& text) { return parse_vcd_text(text, 0); }

Stimulus parse_vcd_file(const std::string& path, int threads) {
  return parse_vcd_text(read_file(path), threads);
}

// ------------------------------------------------------------- writers ----
std::string gv_text(const Netlist& nl, const Library& lib) {
  std::string o;
  o.reserve(nl.insts.size() * 64 + nl.net_names.size() * 12);
  o += "module " + nl.top + " ( );\n";
  std::string ins, wires;
  for (size_t n = 0; n < nl.net_names.size(); ++n) {
    if (nl.const_hidden[n]) continue;
    std::string& dst = nl.net_is_input[n] ? ins : wires;
    dst += dst.empty() ? "" : ",\n    ";
    dst += nl.net_names[n];
  }
  if (!ins.empty()) o += "  input " + ins + ";\n";
  if (!wires.empty()) o += "  wire " + wires + ";\n";
  for (auto& inst : nl.insts) {
    o += "  " + lib.types[inst.type].name + " " + inst.name + " (";
    bool first = true;
    for (auto& pp : inst.in_pins) {
      o += first ? "" : ", ";
      first = false;
      o += "." + pp.first + "(" + nl.net_names[pp.second] + ")";
    }
    for (auto& pp : inst.out_pins) {
      o += first ? "" : ", ";
      first = false;
      o += "." + pp.first + "(" + nl.net_names[pp.second] + ")";
    }
    o += ");\n";
  }
  o += "endmodule\n";
  return o;
}

static std::string fmt_delay(double v) {
  char b[64];
  if (v == (double)(long long)v) std::snprintf(b, sizeof b, "%lld", (long long)v);
  else std::snprintf(b, sizeof b, "%.17g", v);
  return b;
}

std::string sdf_text(const Netlist& nl, const Library& lib) {
  std::string o;
  o += "(DELAYFILE\n  (SDFVERSION \"3.0\")\n  (DESIGN \"" + nl.top + "\")\n  (TIMESCALE 1ps)\n";
  for (auto& inst : nl.insts) {
    o += "  (CELL\n    (CELLTYPE \"" + lib.types[inst.type].name + "\")\n    (INSTANCE " + inst.name +
         ")\n    (DELAY\n      (ABSOLUTE\n";
    for (auto& d : inst.delays) {
      std::string in_pin, out_pin;
      for (auto& pp : inst.in_pins)
        if (pp.second == d.in_net) { in_pin = pp.first; break; }
      for (auto& pp : inst.out_pins)
        if (pp.second == d.out_net) { out_pin = pp.first; break; }
      if (d.out_net <= -2) out_pin = inst.unknown_out[-2 - d.out_net];
      std::string r = fmt_delay(d.rise), f = fmt_delay(d.fall);
      o += "        (IOPATH ";
      if (d.edge == 0) o += "(posedge " + in_pin + ")";
      else if (d.edge == 1) o += "(negedge " + in_pin + ")";
      else o += in_pin;
      o += " " + out_pin + " (" + r + ":" + r + ":" + r + ") (" + f + ":" + f + ":" + f + "))\n";
    }
    o += "      )\n    )\n  )\n";
  }
  o += ")\n";
  return o;
}

static std::string vcd_id(size_t i) {
  std::string s;
  do {
    s += (char)(33 + i % 94);
    i /= 94;
  } while (i);
  return s;
}

std::string vcd_text(const Stimulus& st) {
  std::string o;
  size_t total = 0;
  for (auto& w : st.waves) total += w.size();
  o.reserve(total * 12 + st.names.size() * 32 + 256);
  o += "$date\n  synthetic\n$end\n$version\n  glsgen\n$end\n";
  o += "$timescale " + std::to_string(st.timescale_ps) + "ps $end\n";
  o += "$scope module " + st.scope + " $end\n";
  for (size_t i = 0; i < st.names.size(); ++i)
    o += "$var wire 1 " + vcd_id(i) + " " + st.names[i] + " $end\n";
  o += "$upscope $end\n$enddefinitions $end\n";
  // k-way merge of all signals by time
  std::vector<size_t> pos(st.waves.size(), 0);
  std::vector<std::pair<uint32_t, uint32_t>> heap;  // (time, signal)
  for (size_t i = 0; i < st.waves.size(); ++i)
    if (st.waves[i].size()) heap.push_back({st.waves[i].t[0], (uint32_t)i});
  auto cmp = [](const std::pair<uint32_t, uint32_t>& a, const std::pair<uint32_t, uint32_t>& b) {
    return a > b;
  };
  std::make_heap(heap.begin(), heap.end(), cmp);
  long long last_t = -1;
  static const char vc[4] = {'0', 'x', 'z', '1'};
  char buf[32];
  while (!heap.empty()) {
    std::pop_heap(heap.begin(), heap.end(), cmp);
    auto [t, sig] = heap.back();
    heap.pop_back();
    if ((long long)t != last_t) {
      int k = std::snprintf(buf, sizeof buf, "#%u\n", t);
      o.append(buf, (size_t)k);
      last_t = t;
    }
    o += vc[st.waves[sig].v[pos[sig]] & 3];
    o += vcd_id(sig);
    o += '\n';
    if (++pos[sig] < st.waves[sig].size()) {
      heap.push_back({st.waves[sig].t[pos[sig]], sig});
      std::push_heap(heap.begin(), heap.end(), cmp);
    }
  }
  return o;
}

// ---------------------------------------------------------------- SAIF ----
// Record layout of src/simulate.cpp:85-88,162-163,303-304,373-378 with the
// header of src/simulate.cu:457-481; nets in lexicographic order.
std::string saif_text(std::vector<SaifRecord> recs, uint64_t start, uint64_t end, int timescale_ps,
                      bool with_tc) {
  std::sort(recs.begin(), recs.end(),
            [](const SaifRecord& a, const SaifRecord& b) { return a.name < b.name; });
  std::string o;
  o.reserve(recs.size() * 64 + 128);
  const char* unit = "ps";
  long long res = timescale_ps;
  if (timescale_ps % 1000000 == 0) { unit = "us"; res = timescale_ps / 1000000; }
  else if (timescale_ps % 1000 == 0) { unit = "ns"; res = timescale_ps / 1000; }
  char buf[256];
  std::snprintf(buf, sizeof buf, "(TIMESCALE %lld %s)\n(DURATION %llu)\n(INSTANCE root\n   (NET\n",
                res, unit, (unsigned long long)(end - start));
  o += buf;
  std::vector<const SaifRecord*> hier;
  for (auto& r : recs) {
    size_t lb = r.name.find('[');
    if (lb != std::string::npos && lb + 1 < r.name.size() &&
        !(r.name[lb + 1] >= '0' && r.name[lb + 1] <= '9')) {
      hier.push_back(&r);
      continue;
    }
    o += "      (" + r.name + "\n";
    if (with_tc)
      std::snprintf(buf, sizeof buf, "         (T0 %llu) (T1 %llu) (TX %llu) (TC %llu)\n      )\n",
                    (unsigned long long)r.t0, (unsigned long long)r.t1, (unsigned long long)r.tx,
                    (unsigned long long)r.tc);
    else
      std::snprintf(buf, sizeof buf, "         (T0 %llu) (T1 %llu) (TX %llu)\n      )\n",
                    (unsigned long long)r.t0, (unsigned long long)r.t1, (unsigned long long)r.tx);
    o += buf;
  }
  o += "   )\n\n";
  for (auto* r : hier) {
    size_t lb = r->name.find('['), rb = r->name.find(']');
    std::string inst = r->name.substr(0, lb);
    std::string net = r->name.substr(lb + 1, rb - lb - 1);
    o += "   (INSTANCE " + inst + "\n      (NET\n         (" + net + "\n";
    if (with_tc)
      std::snprintf(buf, sizeof buf,
                    "            (T0 %llu) (T1 %llu) (TX %llu) (TC %llu)\n         )\n      )\n   )\n",
                    (unsigned long long)r->t0, (unsigned long long)r->t1, (unsigned long long)r->tx,
                    (unsigned long long)r->tc);
    else
      std::snprintf(buf, sizeof buf,
                    "            (T0 %llu) (T1 %llu) (TX %llu)\n         )\n      )\n   )\n",
                    (unsigned long long)r->t0, (unsigned long long)r->t1, (unsigned long long)r->tx);
    o += buf;
  }
  o += ")\n";
  return o;
}

}  // namespace gls
