#include "gen.h"

#include <algorithm>
#include <thread>
#include <cmath>
#include <cstring>

namespace gls {

static const double kLn2 = 0.69314718055994530942;

double det_log(double x) {
  if (!(x > 0)) throw GlsError(-1, "det_log of non-positive value");
  // x = m * 2^e with m in [0.5, 1): frexp is exact.
  int e = 0;
  double m = std::frexp(x, &e);
  if (m < 0.70710678118654752440) { m *= 2.0; e -= 1; }
  // ln m = 2 atanh(z), z = (m-1)/(m+1), |z| <= 0.172
  double z = (m - 1.0) / (m + 1.0);
  double z2 = z * z, term = z, sum = 0.0;
  for (int k = 0; k < 30; ++k) {
    sum += term / (double)(2 * k + 1);
    term *= z2;
  }
  return 2.0 * sum + (double)e * kLn2;
}

double det_exp(double x) {
  // x = k ln2 + r, |r| <= ln2/2
  double kf = std::floor(x / kLn2 + 0.5);
  double r = x - kf * kLn2;
  double term = 1.0, sum = 1.0;
  for (int i = 1; i < 30; ++i) {
    term = term * r / (double)i;
    sum += term;
  }
  return std::ldexp(sum, (int)kf);
}

const std::string& generic_vlib_text() {
  static const std::string txt =
      "`timescale 1ps/1ps\n"
      "primitive udp_mux2 (z, i0, i1, s);\n"
      "   output z;\n"
      "   input  i0, i1, s;\n"
      "   table\n"
      "    1  ?  0 :  1;\n"
      "    0  ?  0 :  0;\n"
      "    ?  1  1 :  1;\n"
      "    ?  0  1 :  0;\n"
      "    0  0  x :  0;\n"
      "    1  1  x :  1;\n"
      "   endtable\n"
      "endprimitive\n"
      "`celldefine\n"
      "module GEN_INV_D1 (i,zn);\n  input i;\n  output zn;\n  not (zn, i);\nendmodule\n"
      "module GEN_BUF_D1 (i,z);\n  input i;\n  output z;\n  buf (z, i);\nendmodule\n"
      "module GEN_AND2_D1 (a1,a2,z);\n  input a1;\n  input a2;\n  output z;\n  and (z, a1, a2);\nendmodule\n"
      "module GEN_OR2_D1 (a1,a2,z);\n  input a1;\n  input a2;\n  output z;\n  or (z, a1, a2);\nendmodule\n"
      "module GEN_NAND2_D1 (a1,a2,zn);\n  input a1;\n  input a2;\n  output zn;\n  nand (zn, a1, a2);\nendmodule\n"
      "module GEN_NOR2_D1 (a1,a2,zn);\n  input a1;\n  input a2;\n  output zn;\n  nor (zn, a1, a2);\nendmodule\n"
      "module GEN_XOR2_D1 (a1,a2,z);\n  input a1;\n  input a2;\n  output z;\n  xor (z, a1, a2);\nendmodule\n"
      "module GEN_XNOR2_D1 (a1,a2,zn);\n  input a1;\n  input a2;\n  output zn;\n  xnor (zn, a1, a2);\nendmodule\n"
      "module GEN_MUX2_D1 (i0,i1,s,z);\n  input i0;\n  input i1;\n  input s;\n  output z;\n"
      "  udp_mux2 udpi0 (z, i0, i1, s);\nendmodule\n"
      "module GEN_AOI21_D1 (a1,a2,b,zn);\n  input a1;\n  input a2;\n  input b;\n  output zn;\n"
      "  wire net0, net1;\n  and (net0, a1, a2);\n  or (net1, net0, b);\n  not (zn, net1);\nendmodule\n"
      "module GEN_AOI22_D1 (a1,a2,b1,b2,zn);\n  input a1;\n  input a2;\n  input b1;\n  input b2;\n"
      "  output zn;\n  wire net0, net1, net2;\n  and (net0, a1, a2);\n  and (net1, b1, b2);\n"
      "  or (net2, net0, net1);\n  not (zn, net2);\nendmodule\n"
      "module GEN_OAI22_D1 (a1,a2,b1,b2,zn);\n  input a1;\n  input a2;\n  input b1;\n  input b2;\n"
      "  output zn;\n  wire net0, net1, net2;\n  or (net0, a1, a2);\n  or (net1, b1, b2);\n"
      "  and (net2, net0, net1);\n  not (zn, net2);\nendmodule\n"
      "module GEN_AO221_D1 (a1,a2,b1,b2,c,z);\n  input a1;\n  input a2;\n  input b1;\n  input b2;\n"
      "  input c;\n  output z;\n  wire net0, net1;\n  and (net0, a1, a2);\n  and (net1, b1, b2);\n"
      "  or (z, net0, net1, c);\nendmodule\n"
      "module GEN_AO222_D1 (a1,a2,b1,b2,c1,c2,z);\n  input a1;\n  input a2;\n  input b1;\n  input b2;\n"
      "  input c1;\n  input c2;\n  output z;\n  wire net0, net1, net2;\n  and (net0, c1, c2);\n"
      "  and (net1, a1, a2);\n  and (net2, b1, b2);\n  or (z, net0, net1, net2);\nendmodule\n"
      "`endcelldefine\n";
  return txt;
}

// Generator cell list, in draw order.  `edge_pins` lists the input pins that
// carry separate posedge/negedge IOPATHs (the vlib specify blocks of
// GEN_XOR2_D1, GEN_XNOR2_D1 and GEN_MUX2_D1's select).
struct GenCell {
  const char* name;
  const char* edge_pins;  // comma separated
};
static const GenCell kGenCells[14] = {
    {"GEN_INV_D1", ""},         {"GEN_BUF_D1", ""},
    {"GEN_AND2_D1", ""},        {"GEN_OR2_D1", ""},
    {"GEN_NAND2_D1", ""},       {"GEN_NOR2_D1", ""},
    {"GEN_XOR2_D1", "a1,a2"},   {"GEN_XNOR2_D1", "a1,a2"},
    {"GEN_MUX2_D1", "s"},       {"GEN_AOI21_D1", ""},
    {"GEN_AOI22_D1", ""},       {"GEN_OAI22_D1", ""},
    {"GEN_AO221_D1", ""},       {"GEN_AO222_D1", ""},
};

static bool is_edge_pin(const GenCell& c, const std::string& pin) {
  std::string s = c.edge_pins;
  size_t p = 0;
  while (p <= s.size() && !s.empty()) {
    size_t q = s.find(',', p);
    if (q == std::string::npos) q = s.size();
    if (s.compare(p, q - p, pin) == 0) return true;
    p = q + 1;
  }
  return false;
}

// Approximately standard normal: Irwin-Hall sum of 12 uniforms minus 6.
static double approx_normal(SplitMix64& r) {
  double s = 0;
  for (int i = 0; i < 12; ++i) s += r.u01();
  return s - 6.0;
}

static std::vector<int> level_widths(const NetlistSpec& sp, SplitMix64& rng) {
  int L = sp.n_levels;
  std::vector<int> w(L, 0);
  if (!sp.lognormal_widths) {
    for (int l = 0; l < L; ++l) w[l] = sp.n_cells / L + (l < sp.n_cells % L ? 1 : 0);
    return w;
  }
  // Log-normal raw widths, normalised to n_cells, each >= 1.
  std::vector<double> raw(L);
  double tot = 0;
  for (int l = 0; l < L; ++l) {
    raw[l] = det_exp(sp.lognormal_sigma * approx_normal(rng));
    tot += raw[l];
  }
  long assigned = 0;
  std::vector<std::pair<double, int>> frac;
  for (int l = 0; l < L; ++l) {
    double x = raw[l] / tot * (double)(sp.n_cells - L);
    int base = (int)std::floor(x);
    w[l] = 1 + base;
    assigned += w[l];
    frac.push_back({x - base, l});
  }
  std::sort(frac.begin(), frac.end(), [](auto& a, auto& b) {
    return a.first != b.first ? a.first > b.first : a.second < b.second;
  });
  for (size_t i = 0; assigned < sp.n_cells; ++i, ++assigned) w[frac[i % L].second]++;
  // Organ-pipe order (smallest at both ends, largest in the middle) keeps
  // adjacent levels within a bounded ratio, so every cell can take one input
  // from the level just below it without blowing the fanout cap.
  std::vector<int> sorted = w;
  std::sort(sorted.begin(), sorted.end());
  std::vector<int> out(L);
  int lo = 0, hi = L - 1;
  for (int i = 0; i < L; ++i) {
    if (i % 2 == 0) out[lo++] = sorted[i];
    else out[hi--] = sorted[i];
  }
  return out;
}

Netlist generate_netlist(const NetlistSpec& sp, const Library& lib,
                         std::vector<int>* level_of_cell) {
  if (sp.n_types != 10 && sp.n_types != 12 && sp.n_types != 14)
    throw GlsError(-2, "n_types must be 10, 12 or 14");
  if (sp.n_levels < 1 || sp.n_cells < sp.n_levels || sp.n_pis < 4)
    throw GlsError(-2, "bad netlist spec");
  SplitMix64 rng(sp.seed);
  Netlist nl;
  nl.top = "top";
  std::vector<int> type_ids(sp.n_types);
  for (int i = 0; i < sp.n_types; ++i) {
    type_ids[i] = lib.find(kGenCells[i].name);
    if (type_ids[i] < 0) throw GlsError(-2, std::string("library lacks ") + kGenCells[i].name);
  }
  std::vector<int> widths = level_widths(sp, rng);

  // Nets in creation order: PIs, then one output net per cell, level by
  // level.  level_start[l] = first net index of level l (l = 0 are PIs).
  int total_nets = sp.n_pis + sp.n_cells;
  std::vector<int> level_start(sp.n_levels + 2, 0);
  level_start[0] = 0;
  level_start[1] = sp.n_pis;
  for (int l = 1; l <= sp.n_levels; ++l) level_start[l + 1] = level_start[l] + widths[l - 1];

  for (int i = 0; i < sp.n_pis; ++i) {
    int id = nl.add_net("pi" + std::to_string(i));
    nl.net_is_input[id] = 1;
  }
  for (int c = 0; c < sp.n_cells; ++c) nl.add_net("n" + std::to_string(c));

  // Source weights and cumulative sums for weighted choice.
  std::vector<uint64_t> weight(total_nets, 1);
  if (sp.pareto_fanout) {
    for (int n = 0; n < total_nets; ++n) {
      double w = det_exp(-det_log(rng.u01()) / 1.5);  // Pareto(alpha=1.5), x_m=1
      if (w > 256.0) w = 256.0;
      weight[n] = (uint64_t)std::floor(w);
      if (weight[n] < 1) weight[n] = 1;
    }
  }
  std::vector<uint64_t> cum(total_nets + 1, 0);
  for (int n = 0; n < total_nets; ++n) cum[n + 1] = cum[n] + weight[n];
  std::vector<int> fanout(total_nets, 0);
  const int cap = sp.pareto_fanout ? sp.fanout_cap : (1 << 30);

  auto pick = [&](int lo, int hi) -> int {  // weighted pick in [lo, hi)
    for (int tries = 0; tries < 64; ++tries) {
      uint64_t span = cum[hi] - cum[lo];
      uint64_t r = cum[lo] + rng.next() % span;
      int n = (int)(std::upper_bound(cum.begin() + lo, cum.begin() + hi + 1, r) - cum.begin()) - 1;
      if (fanout[n] < cap) return n;
    }
    for (int n = lo; n < hi; ++n)
      if (fanout[n] < cap) return n;
    return lo + (int)(rng.next() % (uint64_t)(hi - lo));
  };

  nl.insts.reserve(sp.n_cells);
  if (level_of_cell) level_of_cell->assign(sp.n_cells, 0);
  int cell = 0;
  for (int l = 1; l <= sp.n_levels; ++l) {
    for (int i = 0; i < widths[l - 1]; ++i, ++cell) {
      int tsel = (int)(rng.next() % (uint64_t)sp.n_types);
      const CellType& ct = lib.types[type_ids[tsel]];
      int k = (int)ct.inputs.size();
      std::vector<int> ins(k, -1);
      int lead = (int)(rng.next() % (uint64_t)k);
      // one input from the level just below
      ins[lead] = pick(level_start[l - 1], level_start[l]);
      for (int j = 0; j < k; ++j) {
        if (j == lead) continue;
        for (int tries = 0;; ++tries) {
          int n = pick(0, level_start[l]);
          bool dup = false;
          for (int x : ins) dup |= (x == n);
          if (!dup) { ins[j] = n; break; }
          if (tries > 1000) throw GlsError(-2, "generator could not find distinct inputs");
        }
      }
      Instance inst;
      inst.name = "U" + std::to_string(cell);
      inst.type = type_ids[tsel];
      for (int j = 0; j < k; ++j) {
        inst.in_pins.push_back({ct.inputs[j], ins[j]});
        fanout[ins[j]]++;
      }
      int out = sp.n_pis + cell;
      inst.out_pins.push_back({ct.outputs[0], out});
      for (int j = 0; j < k; ++j) {
        auto draw = [&]() { return (double)rng.range(sp.delay_min, sp.delay_max); };
        if (is_edge_pin(kGenCells[tsel], ct.inputs[j])) {
          for (int e = 0; e < 2; ++e) {
            IoPath d;
            d.edge = e;
            d.in_net = ins[j];
            d.out_net = out;
            d.rise = draw();
            d.fall = draw();
            inst.delays.push_back(d);
          }
        } else {
          IoPath d;
          d.edge = 2;
          d.in_net = ins[j];
          d.out_net = out;
          d.rise = draw();
          d.fall = draw();
          inst.delays.push_back(d);
        }
      }
      if (level_of_cell) (*level_of_cell)[cell] = l;
      nl.insts.push_back(std::move(inst));
    }
  }
  return nl;
}

namespace {

// The stimulus walk draws from one SplitMix64 stream in a fixed order.  A
// draw is a function of its index alone (the generator's state is a
// counter), so blocks of draws and the logarithms of their uniforms are
// computed ahead on several threads, and the serial walk only consumes
// them: the result is the serial pass's, bit for bit, on any thread count.
class AheadStream {
 public:
  AheadStream(uint64_t seed, int threads) : seed_(seed), threads_(std::max(1, threads)) {}
  uint64_t next() {
    ensure();
    return raw_[idx_++ - base_];
  }
  double u01() { return to_u01(next()); }
  // det_log(u01()) of the next draw
  double log_u01() {
    ensure();
    return logu_[idx_++ - base_];
  }

 private:
  static double to_u01(uint64_t r) { return ((r >> 11) + 1) * (1.0 / 9007199254740992.0); }
  void ensure() {
    if (idx_ >= base_ && idx_ < base_ + raw_.size()) return;
    base_ = idx_;
    const size_t B = raw_.empty() ? (1u << 16) : std::min<size_t>(raw_.size() * 4, 1u << 22);
    raw_.resize(B);
    logu_.resize(B);
    auto work = [&](size_t a, size_t b) {
      for (size_t i = a; i < b; ++i) {
        // the (base + i + 1)-th draw of a SplitMix64 seeded with seed_
        uint64_t z = seed_ + (base_ + i + 1) * 0x9E3779B97F4A7C15ull;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        raw_[i] = z ^ (z >> 31);
        logu_[i] = det_log(to_u01(raw_[i]));
      }
    };
    const int T = B >= (1u << 18) ? threads_ : 1;
    if (T == 1) {
      work(0, B);
      return;
    }
    std::vector<std::thread> th;
    for (int t = 0; t < T; ++t) th.emplace_back(work, B * t / T, B * (t + 1) / T);
    for (auto& x : th) x.join();
  }
  uint64_t seed_, idx_ = 0, base_ = 0;
  int threads_;
  std::vector<uint64_t> raw_;
  std::vector<double> logu_;
};

}  // namespace

Stimulus generate_stimulus(const Netlist& nl, const StimulusSpec& sp) {
  AheadStream rng(sp.seed ^ 0xA5A5A5A5DEADBEEFull, host_threads(0));
  Stimulus st;
  st.scope = nl.top;
  for (size_t n = 0; n < nl.net_names.size(); ++n) {
    if (!nl.net_is_input[n]) continue;
    st.names.push_back(nl.net_names[n]);
    Wave w;
    uint8_t cur = (rng.next() & 1) ? V1 : V0;
    w.push(0, cur);
    double t = 0;
    for (;;) {
      double gap = -sp.mean_gap * rng.log_u01();
      uint64_t g = (uint64_t)std::floor(gap + 0.5);
      if (g < 1) g = 1;
      t += (double)g;
      if (t > (double)sp.t_end) break;
      uint8_t nv;
      if (cur == VX) nv = (rng.next() & 1) ? V1 : V0;
      else if (rng.u01() <= sp.x_prob) nv = VX;
      else nv = (cur == V0) ? V1 : V0;
      cur = nv;
      w.push((uint32_t)t, cur);
    }
    st.waves.push_back(std::move(w));
  }
  return st;
}

NetlistSpec config_netlist_spec(int ci) {
  NetlistSpec s;
  s.seed = 0x5EED0000ull + (uint64_t)ci;
  switch (ci) {
    case 0:
      s.n_cells = 2000; s.n_pis = 64; s.n_levels = 20; s.n_types = 10;
      break;
    case 1:
      s.n_cells = 200000; s.n_pis = 1024; s.n_levels = 40; s.n_types = 12;
      s.pareto_fanout = true;
      break;
    case 2:
      s.n_cells = 2000000; s.n_pis = 4096; s.n_levels = 60; s.n_types = 12;
      s.pareto_fanout = true; s.lognormal_widths = true;
      break;
    case 3:  // the 200k-cell shape with 5- and 6-input cells in the mix
      s.n_cells = 200000; s.n_pis = 1024; s.n_levels = 40; s.n_types = 14;
      s.pareto_fanout = true;
      break;
    default:
      throw GlsError(-2, "unknown config index");
  }
  return s;
}

StimulusSpec config_stimulus_spec(int ci) {
  StimulusSpec s;
  s.seed = 0x5EED0000ull + (uint64_t)ci;
  switch (ci) {
    case 0: s.t_end = 2000000; s.mean_gap = 1000; break;
    case 1: s.t_end = 20000000; s.mean_gap = 500; break;
    case 2: s.t_end = 20000000; s.mean_gap = 200; break;
    case 3: s.t_end = 20000000; s.mean_gap = 500; break;
    default: throw GlsError(-2, "unknown config index");
  }
  return s;
}

}  // namespace gls
