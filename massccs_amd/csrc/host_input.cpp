// input.json reading with the reference's keys and defaults
// (src/Input.cpp:49-263, src/headers/Constants.h:68-83).
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iterator>
#include <map>
#include <memory>
#include <thread>

#include "host.hpp"

namespace mccs {
namespace {

// Minimal JSON value: enough for the flat object input.json is.
struct JVal {
  enum Kind { NUL, BOOL, NUM, STR, ARR, OBJ } kind = NUL;
  bool b = false;
  double num = 0;
  bool is_int = false;
  std::string raw_num, str;
  std::vector<JVal> arr;
  std::vector<std::pair<std::string, JVal>> obj;
};

class JParser {
 public:
  explicit JParser(const std::string &s) : s_(s) {}
  JVal parse() {
    JVal v = value();
    ws();
    if (i_ != s_.size()) fail("trailing characters");
    return v;
  }

 private:
  const std::string &s_;
  size_t i_ = 0;
  [[noreturn]] void fail(const char *what) {
    throw Error(MCCS_ERR_PARSE, std::string("input.json: ") + what + " at offset " +
                                    std::to_string(i_));
  }
  void ws() {
    while (i_ < s_.size() && std::isspace((unsigned char)s_[i_])) i_++;
  }
  JVal value() {
    ws();
    if (i_ >= s_.size()) fail("unexpected end");
    char c = s_[i_];
    if (c == '{') return object();
    if (c == '[') return array();
    if (c == '"') {
      JVal v;
      v.kind = JVal::STR;
      v.str = string();
      return v;
    }
    if (c == 't' && s_.compare(i_, 4, "true") == 0) { i_ += 4; JVal v; v.kind = JVal::BOOL; v.b = true; return v; }
    if (c == 'f' && s_.compare(i_, 5, "false") == 0) { i_ += 5; JVal v; v.kind = JVal::BOOL; return v; }
    if (c == 'n' && s_.compare(i_, 4, "null") == 0) { i_ += 4; return JVal(); }
    return number();
  }
  std::string string() {
    std::string out;
    i_++;  // opening quote
    while (i_ < s_.size() && s_[i_] != '"') {
      char c = s_[i_++];
      if (c == '\\') {
        if (i_ >= s_.size()) fail("bad escape");
        char e = s_[i_++];
        switch (e) {
          case 'n': out += '\n'; break;
          case 't': out += '\t'; break;
          case 'r': out += '\r'; break;
          case 'b': out += '\b'; break;
          case 'f': out += '\f'; break;
          case 'u': {
            if (i_ + 4 > s_.size()) fail("bad unicode escape");
            unsigned cp = std::stoul(s_.substr(i_, 4), nullptr, 16);
            i_ += 4;
            if (cp < 0x80) out += (char)cp;
            else if (cp < 0x800) { out += (char)(0xC0 | (cp >> 6)); out += (char)(0x80 | (cp & 0x3F)); }
            else { out += (char)(0xE0 | (cp >> 12)); out += (char)(0x80 | ((cp >> 6) & 0x3F)); out += (char)(0x80 | (cp & 0x3F)); }
            break;
          }
          default: out += e;
        }
      } else {
        out += c;
      }
    }
    if (i_ >= s_.size()) fail("unterminated string");
    i_++;
    return out;
  }
  JVal number() {
    size_t b = i_;
    if (s_[i_] == '-') i_++;
    bool frac = false;
    while (i_ < s_.size() && (std::isdigit((unsigned char)s_[i_]) || s_[i_] == '.' ||
                              s_[i_] == 'e' || s_[i_] == 'E' || s_[i_] == '+' || s_[i_] == '-')) {
      if (s_[i_] == '.' || s_[i_] == 'e' || s_[i_] == 'E') frac = true;
      i_++;
    }
    if (b == i_) fail("unexpected character");
    JVal v;
    v.kind = JVal::NUM;
    v.raw_num = s_.substr(b, i_ - b);
    char *end = nullptr;
    v.num = std::strtod(v.raw_num.c_str(), &end);
    if (!end || *end) fail("bad number");
    v.is_int = !frac;
    return v;
  }
  JVal array() {
    JVal v;
    v.kind = JVal::ARR;
    i_++;
    ws();
    if (i_ < s_.size() && s_[i_] == ']') { i_++; return v; }
    for (;;) {
      v.arr.push_back(value());
      ws();
      if (i_ >= s_.size()) fail("unterminated array");
      if (s_[i_] == ',') { i_++; continue; }
      if (s_[i_] == ']') { i_++; return v; }
      fail("expected , or ]");
    }
  }
  JVal object() {
    JVal v;
    v.kind = JVal::OBJ;
    i_++;
    ws();
    if (i_ < s_.size() && s_[i_] == '}') { i_++; return v; }
    for (;;) {
      ws();
      if (i_ >= s_.size() || s_[i_] != '"') fail("expected key");
      std::string k = string();
      ws();
      if (i_ >= s_.size() || s_[i_] != ':') fail("expected :");
      i_++;
      v.obj.emplace_back(k, value());
      ws();
      if (i_ >= s_.size()) fail("unterminated object");
      if (s_[i_] == ',') { i_++; continue; }
      if (s_[i_] == '}') { i_++; return v; }
      fail("expected , or }");
    }
  }
};

const JVal *member(const JVal &o, const char *key) {
  for (auto &kv : o.obj)
    if (kv.first == key) return &kv.second;
  return nullptr;
}

uint32_t get_uint(const JVal &v, const char *key) {
  if (v.kind != JVal::NUM || !v.is_int || v.num < 0 || v.num > 4294967295.0)
    throw Error(MCCS_ERR_PARSE, std::string("input.json: ") + key + " must be an unsigned integer");
  return (uint32_t)v.num;
}

double get_double(const JVal &v, const char *key) {
  if (v.kind != JVal::NUM)
    throw Error(MCCS_ERR_PARSE, std::string("input.json: ") + key + " must be a number");
  return v.num;
}

std::string get_string(const JVal &v, const char *key) {
  if (v.kind != JVal::STR)
    throw Error(MCCS_ERR_PARSE, std::string("input.json: ") + key + " format error");
  return v.str;
}

int yes_no(const JVal &o, const char *key, int dflt, const char *msg) {
  const JVal *v = member(o, key);
  if (!v) return dflt;
  std::string s = get_string(*v, key);
  if (s == "yes") return 1;
  if (s == "no") return 0;
  throw Error(MCCS_ERR_INVALID, msg);
}

}  // namespace

std::string gas_name(int flag) {
  switch (flag) {
    case MCCS_GAS_HE: return "He";
    case MCCS_GAS_N2: return "N2";
    case MCCS_GAS_CO2: return "CO2";
    case MCCS_GAS_AR: return "Ar";
    case MCCS_GAS_CO2_1SITE: return "co2";
  }
  return "?";
}

mccs_config default_config() {
  mccs_config c;
  std::memset(&c, 0, sizeof c);
  c.n_probe = 10000;        // NPROBE
  c.n_iter = 10;            // NITER
  c.seed = 20162104;        // SEED
  unsigned hc = std::thread::hardware_concurrency();
  c.nthreads = hc ? hc : 1; // omp_get_max_threads() in the reference
  c.dt = 10.0;              // TIMESTEP
  c.temperature = 298.0;    // TEMPERATURE
  c.skin = 0.01;            // SKIN
  c.lj_cutoff = 12.0;       // SHORT_CUTOFF
  c.coul_cutoff = 25.0;     // LONG_CUTOFF
  c.gas = MCCS_GAS_HE;
  c.short_range_cutoff = 1;
  c.long_range_flag = 0;
  c.long_range_cutoff = 1;
  c.polarizability = 1;     // yes for He (src/Input.cpp:232-240)
  c.alpha = 0.204956;       // ALPHA_HE
  c.alpha_set = 1;
  return c;
}

// Input::readInputFile, src/Input.cpp:49-263.
void load_json_config(const std::string &path, mccs_config &c, std::string &gas_str) {
  std::ifstream in(path);
  if (!in) throw Error(MCCS_ERR_IO, "cannot open input file " + path);
  std::string text((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
  JVal d = JParser(text).parse();
  if (d.kind != JVal::OBJ) throw Error(MCCS_ERR_PARSE, "input.json: top level must be an object");
  c = default_config();
  const JVal *v = member(d, "targetFileName");
  if (!v) throw Error(MCCS_ERR_INVALID, "Error input.json: targetFileName not found");
  std::string tf = get_string(*v, "targetFileName");
  if (tf.size() >= sizeof c.target_file) throw Error(MCCS_ERR_INVALID, "targetFileName too long");
  std::strcpy(c.target_file, tf.c_str());
  if ((v = member(d, "numberProbe"))) c.n_probe = get_uint(*v, "numberProbe");
  if ((v = member(d, "nIter"))) c.n_iter = get_uint(*v, "nIter");
  if ((v = member(d, "seed"))) c.seed = get_uint(*v, "seed");
  if ((v = member(d, "nthreads"))) c.nthreads = get_uint(*v, "nthreads");
  if ((v = member(d, "dt"))) c.dt = get_double(*v, "dt");
  if ((v = member(d, "Temp"))) c.temperature = get_double(*v, "Temp");
  if ((v = member(d, "skin"))) c.skin = get_double(*v, "skin");
  gas_str = "He";
  c.gas = MCCS_GAS_HE;
  if ((v = member(d, "GasBuffer"))) {
    gas_str = get_string(*v, "GasBuffer");
    if (gas_str == "He") c.gas = MCCS_GAS_HE;
    else if (gas_str == "N2") c.gas = MCCS_GAS_N2;
    else if (gas_str == "CO2") c.gas = MCCS_GAS_CO2;
    else if (gas_str == "Ar") c.gas = MCCS_GAS_AR;
    else if (gas_str == "co2") c.gas = MCCS_GAS_CO2_1SITE;
    else throw Error(MCCS_ERR_INVALID, "only available the follow gas buffer types: He Ar N2 CO2");
  }
  c.equipotential = yes_no(d, "Equipotential", 0, "need to choice Equipotential: yes or no");
  // not a reference key: the corrected chi reference of the CO2 probe
  c.co2_vcm_fix = yes_no(d, "CO2 vcm fix", 0, "need to choice CO2 vcm fix: yes or no");
  c.short_range_cutoff = yes_no(d, "Short-range cutoff", 1, "need to choice Short-range cutoff: yes or no");
  if ((v = member(d, "LJ-cutoff"))) c.lj_cutoff = get_double(*v, "LJ-cutoff");
  c.long_range_flag = yes_no(d, "Long-range forces", 0, "need specify only yes or no for coulomb interactions");
  c.long_range_cutoff = yes_no(d, "Long-range cutoff", 1, "need to choice Long-range cutoff: yes or no");
  if ((v = member(d, "Coul-cutoff"))) c.coul_cutoff = get_double(*v, "Coul-cutoff");
  c.polarizability = yes_no(d, "polarizability", c.gas == MCCS_GAS_HE ? 1 : 0,
                            "need to choice polarizability: yes or no");
  c.alpha_set = 0;
  c.alpha = 0.0;
  if ((v = member(d, "alpha"))) {
    c.alpha = get_double(*v, "alpha");
    c.alpha_set = 1;
  } else if (c.polarizability == 1) {
    switch (c.gas) {
      case MCCS_GAS_HE: c.alpha = 0.204956; break;   // ALPHA_HE
      case MCCS_GAS_N2: c.alpha = 1.710; break;      // ALPHA_N2
      case MCCS_GAS_CO2: c.alpha = 2.911; break;     // ALPHA_CO2
      case MCCS_GAS_AR: c.alpha = 1.6411; break;     // ALPHA_AR
      case MCCS_GAS_CO2_1SITE: c.alpha = 2.911; break;
    }
    c.alpha_set = 1;
  }
  c.user_ff = 0;
  if ((v = member(d, "force-field"))) {
    std::string ff = get_string(*v, "force-field");
    if (ff.size() >= sizeof c.force_field) throw Error(MCCS_ERR_INVALID, "force-field too long");
    std::strcpy(c.force_field, ff.c_str());
    c.user_ff = 1;
  }
}

// Input::printReadInput, src/Input.cpp:313-345.
std::string format_config(const mccs_config &c, const std::string &gas_str) {
  std::ostringstream o;
  auto yn = [](int f) { return f ? "yes" : "no"; };
  o << "*********************************************************\n";
  o << "INPUT:: Simulation Parameters \n";
  o << "*********************************************************\n";
  o << "target filename                  : " << c.target_file << "\n";
  o << "number of probe                  : " << c.n_probe << "\n";
  o << "number of iterarions             : " << c.n_iter << "\n";
  o << "number of threads                : " << c.nthreads << "\n";
  o << "seed number                      : " << c.seed << "\n";
  o << "gas buffer                       : " << gas_str << "\n";
  o << "Target Temperature (K)           : " << c.temperature << "\n";
  o << "timestep (fs)                    : " << c.dt << "\n";
  o << "Skin cell size (Ang)             : " << c.skin << "\n";
  o << "Equipotential                    : " << yn(c.equipotential) << "\n";
  o << "Cut short-range interaction      : " << yn(c.short_range_cutoff) << "\n";
  if (c.short_range_cutoff == 1) o << "LJ cutoff (Ang)                  : " << c.lj_cutoff << "\n";
  o << "Apply long-range interaction     : " << yn(c.long_range_flag) << "\n";
  o << "Cut long-range interaction       : " << yn(c.long_range_cutoff) << "\n";
  if (c.long_range_cutoff == 1) o << "Coulomb cutoff (Ang)             : " << c.coul_cutoff << "\n";
  o << "Apply induced-dipole interaction : " << yn(c.polarizability) << "\n";
  if (c.polarizability == 1) o << "alpha (Ang^3)                    : " << c.alpha << "\n";
  if (c.user_ff == 1) o << "force-field                      : " << c.force_field << "\n";
  return o.str();
}

// Force-type selection, src/System.cpp:49-87.
int select_force_type(const mccs_config &c) {
  const int s = c.short_range_cutoff, l = c.long_range_flag, lc = c.long_range_cutoff,
            p = c.polarizability;
  if (c.gas == MCCS_GAS_HE || c.gas == MCCS_GAS_AR || c.gas == MCCS_GAS_CO2_1SITE) {
    if (s == 0 && l == 0) return 1;
    if (s == 1 && l == 0) return 2;
    if (s == 0 && p == 1 && lc == 0) return 3;
    if (s == 1 && p == 1 && lc == 1) return 4;
    return 2;
  }
  if (s == 0 && l == 0 && p == 0) return 1;
  if (s == 1 && l == 0 && p == 0) return 2;
  if (s == 0 && l == 1 && lc == 0 && p == 0) return 3;
  if (s == 1 && l == 1 && lc == 1 && p == 0) return 4;
  if (s == 0 && l == 1 && lc == 0 && p == 1) return 5;
  if (s == 1 && l == 1 && lc == 1 && p == 1) return 6;
  return 2;
}

}  // namespace mccs
