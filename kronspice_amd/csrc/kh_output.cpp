// kh_output.cpp — the reference's result formats written from what the
// batched analyses return, byte for byte as the reference's tools write them:
//
//  * the JSON document of `kronspice -json` (output/json.go:18-316): a
//    Document with one analysis, as encoding/json's Encoder writes it with
//    SetIndent("", "  ") — struct fields in declaration order, omitempty
//    fields left out when empty, map keys sorted, floats in the shortest form
//    that round-trips ('f' notation for 1e-6 <= |v| < 1e21, else 'e' with
//    the exponent's leading zero dropped), strings escaped HTML-safe, a
//    newline after the document;
//  * the operating point's "name,value" CSV (output/csv.go:15-48) as
//    encoding/csv writes it: fields quoted when they hold a comma, a quote,
//    a line break or start with a space, numbers as strconv.FormatFloat(v,
//    'g', 12, 64).
//
// Host code only; it reads host arrays the caller copied back.
#include <algorithm>
#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "../../include/kronhip.h"

namespace {

// ---- a JSON value tree, serialized as encoding/json + Indent ------------------------------
struct J {
  enum T { NUL, BOOL, NUM, STR, ARR, OBJ } t = NUL;
  bool b = false;
  double n = 0;
  std::string s;
  std::vector<J> a;                              // ARR
  std::vector<std::pair<std::string, J>> o;      // OBJ, in emission order
  static J null() { return J(); }
  static J boolean(bool v) { J j; j.t = BOOL; j.b = v; return j; }
  static J num(double v) { J j; j.t = NUM; j.n = v; return j; }
  static J str(std::string v) { J j; j.t = STR; j.s = std::move(v); return j; }
  static J arr() { J j; j.t = ARR; return j; }
  static J obj() { J j; j.t = OBJ; return j; }
  J& add(const std::string& k, J v) { o.emplace_back(k, std::move(v)); return o.back().second; }
  J& push(J v) { a.push_back(std::move(v)); return a.back(); }
};

// encoding/json's float encoding (float64): shortest round-trip digits.
void put_float(std::string& out, double v) {
  char buf[400];
  const double av = std::fabs(v);
  const bool e = av != 0 && (av < 1e-6 || av >= 1e21);
  auto r = std::to_chars(buf, buf + sizeof(buf), v, e ? std::chars_format::scientific : std::chars_format::fixed);
  std::string t(buf, r.ptr);
  if (e) {  // e-07 -> e-7
    const size_t n = t.size();
    if (n >= 4 && t[n - 4] == 'e' && t[n - 3] == '-' && t[n - 2] == '0') t.erase(n - 2, 1);
  }
  out += t;
}

// encoding/json's string encoding with HTML escaping (the Encoder's default).
void put_string(std::string& out, const std::string& s) {
  static const char* hex = "0123456789abcdef";
  out += '"';
  for (size_t i = 0; i < s.size();) {
    const unsigned char c = (unsigned char)s[i];
    if (c < 0x80) {
      if (c >= 0x20 && c != '"' && c != '\\' && c != '<' && c != '>' && c != '&') {
        out += (char)c;
      } else {
        switch (c) {
          case '\\': out += "\\\\"; break;
          case '"': out += "\\\""; break;
          case '\b': out += "\\b"; break;
          case '\f': out += "\\f"; break;
          case '\n': out += "\\n"; break;
          case '\r': out += "\\r"; break;
          case '\t': out += "\\t"; break;
          default:
            out += "\\u00";
            out += hex[c >> 4];
            out += hex[c & 15];
        }
      }
      i++;
      continue;
    }
    // a UTF-8 sequence: copied when valid; U+2028 and U+2029 escaped; an
    // invalid byte becomes �
    int len = (c >> 5) == 6 ? 2 : (c >> 4) == 14 ? 3 : (c >> 3) == 30 ? 4 : 0;
    bool ok = len > 0 && i + len <= s.size();
    uint32_t cp = 0;
    if (ok) {
      cp = c & (0x7f >> len);
      for (int k = 1; k < len; k++) {
        const unsigned char d = (unsigned char)s[i + k];
        if ((d >> 6) != 2) { ok = false; break; }
        cp = (cp << 6) | (d & 0x3f);
      }
      const uint32_t lo[5] = {0, 0, 0x80, 0x800, 0x10000};
      if (ok && (cp < lo[len] || cp > 0x10ffff || (cp >= 0xd800 && cp <= 0xdfff))) ok = false;
    }
    if (!ok) {
      out += "\\ufffd";
      i++;
      continue;
    }
    if (cp == 0x2028 || cp == 0x2029) {
      out += cp == 0x2028 ? "\\u2028" : "\\u2029";
    } else {
      out.append(s, i, len);
    }
    i += len;
  }
  out += '"';
}

void put(std::string& out, const J& j, int depth) {
  auto nl = [&](int d) {
    out += '\n';
    for (int k = 0; k < d; k++) out += "  ";
  };
  switch (j.t) {
    case J::NUL: out += "null"; break;
    case J::BOOL: out += j.b ? "true" : "false"; break;
    case J::NUM: put_float(out, j.n); break;
    case J::STR: put_string(out, j.s); break;
    case J::ARR:
      if (j.a.empty()) { out += "[]"; break; }
      out += '[';
      for (size_t k = 0; k < j.a.size(); k++) {
        nl(depth + 1);
        put(out, j.a[k], depth + 1);
        if (k + 1 < j.a.size()) out += ',';
      }
      nl(depth);
      out += ']';
      break;
    case J::OBJ:
      if (j.o.empty()) { out += "{}"; break; }
      out += '{';
      for (size_t k = 0; k < j.o.size(); k++) {
        nl(depth + 1);
        put_string(out, j.o[k].first);
        out += ": ";
        put(out, j.o[k].second, depth + 1);
        if (k + 1 < j.o.size()) out += ',';
      }
      nl(depth);
      out += '}';
      break;
  }
}

// ---- encoding/csv ------------------------------------------------------------------------
void csv_field(std::string& out, const std::string& f) {
  bool q = false;
  if (f == "\\.") q = true;
  for (char c : f)
    if (c == ',' || c == '"' || c == '\r' || c == '\n') q = true;
  if (!f.empty() && (f[0] == ' ' || f[0] == '\t' || f[0] == '\v' || f[0] == '\f')) q = true;
  if (!q) {
    out += f;
    return;
  }
  out += '"';
  for (char c : f) {
    if (c == '"') out += "\"\"";
    else out += c;
  }
  out += '"';
}
void csv_row(std::string& out, const std::string& a, const std::string& b) {
  csv_field(out, a);
  out += ',';
  csv_field(out, b);
  out += '\n';
}

// strconv.FormatFloat(v, 'g', prec, 64): C's %.{prec}g chooses the same
// notation and digits (exponent of at least two digits); Go spells the
// special values NaN, +Inf and -Inf.
std::string go_g(double v, int prec) {
  if (std::isnan(v)) return "NaN";
  if (std::isinf(v)) return v > 0 ? "+Inf" : "-Inf";
  char buf[64];
  std::snprintf(buf, sizeof(buf), "%.*g", prec, v);
  return buf;
}

std::string lower(std::string s) {
  for (auto& c : s)
    if (c >= 'A' && c <= 'Z') c = (char)(c - 'A' + 'a');
  return s;
}

// circuit.Kind.String (circuit/circuit.go:109-140)
const char* kind_name(int k) {
  switch (k) {
    case KH_ELEM_R: return "resistor";
    case KH_ELEM_C: return "capacitor";
    case KH_ELEM_L: return "inductor";
    case KH_ELEM_V: return "voltage source";
    case KH_ELEM_I: return "current source";
    case KH_ELEM_D: return "diode";
    case KH_ELEM_M: return "MOSFET";
    case KH_ELEM_Q: return "BJT";
    case KH_ELEM_E: return "VCVS";
    case KH_ELEM_F: return "CCCS";
    case KH_ELEM_G: return "VCCS";
    case KH_ELEM_H: return "CCVS";
    case KH_ELEM_K: return "K";
  }
  return "unknown";
}

// The element's modeled ports in channel order, with their notes (the
// SetIntrinsic calls of device/passive.go:51-59, device/reactive.go:45, :216,
// device/source.go:221-234, device/diode.go:125-126, device/mosfet.go:211-212,
// device/bjt.go:238-239, device/controlled.go:49-50): (channel, note).
std::vector<std::pair<int, const char*>> ports_of(int kind, bool negative) {
  const char* inside = "terminals connect inside the network";
  switch (kind) {
    case KH_ELEM_R:
      if (negative) return {{0, "idealized negative resistance: delivers v²/|R| as model-supplied power"}};
      return {{0, inside}, {2, "Joule heating i²R to surroundings"}};
    case KH_ELEM_C: return {{0, "terminals connect inside the network; stores ½Cv²"}};
    case KH_ELEM_L: return {{0, "terminals connect inside the network; stores ½Li²"}};
    case KH_ELEM_V:
    case KH_ELEM_I: return {{0, "impressed source: exchanges energy with the exterior"}};
    case KH_ELEM_D:
      return {{0, "terminals connect inside the network; stores junction charge"},
              {2, "junction and series-resistance heating to surroundings"}};
    case KH_ELEM_M:
      return {{0, "terminals connect inside the network; stores gate and junction charge"},
              {2, "channel, junction and series-resistance heating to surroundings"}};
    case KH_ELEM_Q:
      return {{0, "terminals connect inside the network; stores junction charge"},
              {2, "junction, transport and series-resistance heating to surroundings"}};
    case KH_ELEM_E:
    case KH_ELEM_F:
    case KH_ELEM_G:
    case KH_ELEM_H: return {{0, "idealized controlled source: its power, of either sign, is model-supplied"}};
    case KH_ELEM_K: return {{0, "couples two inductors inside the network; stores M·i1·i2"}};  // device/mutual.go:46-47
  }
  return {};
}
const char* channel_name(int ch) { return ch == 0 ? "electrical" : "thermal"; }

// The failure classes of cmd/kronspice/main.go:284-305 for the library's
// per-circuit statuses.
const char* error_class(kh_status s) {
  switch (s) {
    case KH_LEDGER: return "ledger";
    case KH_NOCONV:
    case KH_TSTEP_TOO_SMALL:
    case KH_TOO_MANY_STEPS: return "no-convergence";
    case KH_SINGULAR:
    case KH_NONFINITE: return "singular";
    case KH_INTERRUPTED: return "timeout";
    case KH_BAD_TRAN_SPEC: return "invalid";
  }
  return "error";
}

bool names_ok(const kh_names* nm) {
  if (!nm || nm->num_nodes < 1 || nm->n_elems < 0 || !nm->node_names) return false;
  if (nm->n_elems > 0 && (!nm->elem_names || !nm->elems)) return false;
  for (int i = 0; i < nm->num_nodes; i++)
    if (!nm->node_names[i]) return false;
  for (int k = 0; k < nm->n_elems; k++) {
    if (!nm->elem_names[k]) return false;
    for (int t = 0; t < 2; t++)
      if (nm->elems[k].node[t] < 0 || nm->elems[k].node[t] >= nm->num_nodes) return false;
  }
  return true;
}

int64_t emit(const std::string& s, char* buf, int64_t cap) {
  if (buf && cap > 0) std::memcpy(buf, s.data(), (size_t)std::min<int64_t>(cap, (int64_t)s.size()));
  return (int64_t)s.size();
}

// The document around one analysis, or the failed run's (output.Document).
J document(const kh_names* nm, const char* version, kh_status status, J* analysis, const char* kind) {
  J d = J::obj();
  d.add("tool", J::str("kronspice"));
  d.add("version", J::str(version && *version ? version : "dev"));
  d.add("ok", J::boolean(status == KH_OK));
  if (status != KH_OK) {
    J e = J::obj();
    e.add("class", J::str(error_class(status)));
    e.add("message", J::str(kh_status_string(status)));
    d.add("error", std::move(e));
  }
  if (nm->title && *nm->title) d.add("title", J::str(nm->title));
  J as = J::arr();
  if (status == KH_OK && analysis) {
    J a = J::obj();
    a.add("kind", J::str(kind));
    a.add(kind, std::move(*analysis));
    as.push(std::move(a));
  }
  d.add("analyses", std::move(as));
  return d;
}

// Node names sorted (Go map keys), each with its value from f(node index).
template <class F>
J node_map(const kh_names* nm, F f) {
  std::vector<std::pair<std::string, int>> keys;
  for (int i = 1; i < nm->num_nodes; i++) keys.emplace_back(nm->node_names[i], i);
  // Go sorts map keys by their bytes, unsigned
  std::sort(keys.begin(), keys.end(), [](const auto& a, const auto& b) {
    const int c = std::memcmp(a.first.data(), b.first.data(), std::min(a.first.size(), b.first.size()));
    return c != 0 ? c < 0 : a.first.size() < b.first.size();
  });
  J m = J::obj();
  for (auto& k : keys) m.add(k.first, f(k.second));
  return m;
}

}  // namespace

extern "C" {

int64_t kh_format_op(int32_t format, const kh_names* nm, const char* version, kh_status status, const double* x,
                     const double* report, const double* ledger, char* buf, int64_t cap) {
  if (!names_ok(nm) || (format != KH_FORMAT_JSON && format != KH_FORMAT_CSV)) return -1;
  const bool have = status == KH_OK;
  if (have && (!x || (nm->n_elems > 0 && !report) || !ledger)) return -1;
  const int nel = nm->n_elems;
  auto rp = [&](int k, int f) { return report[(int64_t)k * KH_REPORT_FIELDS + f]; };
  auto negative = [&](int k) {  // a negative resistance supplies power and has no thermal port
    return nm->elems[k].kind == KH_ELEM_R && rp(k, KH_RP_CHANNEL) < 0;
  };
  if (format == KH_FORMAT_CSV) {
    if (!have) return emit("", buf, cap);
    std::string out;
    csv_row(out, "name", "value");
    for (int i = 1; i < nm->num_nodes; i++)
      csv_row(out, std::string("v(") + nm->node_names[i] + ")", go_g(x[i - 1], 12));
    for (int k = 0; k < nel; k++) {
      const std::string n = lower(nm->elem_names[k]);
      csv_row(out, "i(" + n + ")", go_g(rp(k, KH_RP_CURRENT), 12));
      csv_row(out, "p(" + n + ")", go_g(rp(k, KH_RP_POWER), 12));
    }
    for (int k = 0; k < nel; k++) {
      const int ch_ext = (int)rp(k, KH_RP_CHANNEL);
      for (auto& pt : ports_of(nm->elems[k].kind, negative(k))) {
        const double pw = pt.first == ch_ext ? rp(k, KH_RP_EXTERIOR) : 0.0;
        csv_row(out, "ledger(" + lower(nm->elem_names[k]) + "," + channel_name(pt.first) + ")", go_g(pw, 12));
      }
    }
    csv_row(out, "ledger.inflow", go_g(ledger[KH_LG_INFLOW], 12));
    csv_row(out, "ledger.outflow", go_g(ledger[KH_LG_OUTFLOW], 12));
    csv_row(out, "ledger.dissipation", go_g(ledger[KH_LG_DISSIPATION], 12));
    csv_row(out, "ledger.residual", go_g(ledger[KH_LG_RESIDUAL], 12));
    return emit(out, buf, cap);
  }
  J opdoc;
  if (have) {
    opdoc = J::obj();
    opdoc.add("nodes", node_map(nm, [&](int i) { return J::num(x[i - 1]); }));
    J els = nel > 0 ? J::arr() : J::null();
    for (int k = 0; k < nel; k++) {
      J e = J::obj();
      e.add("name", J::str(nm->elem_names[k]));
      e.add("kind", J::str(kind_name(nm->elems[k].kind)));
      J nodes = J::arr();
      // a K element's nodes are ground (circuit/builder.go:80-84); node[] holds the inductors
      const bool kk = nm->elems[k].kind == KH_ELEM_K;
      nodes.push(J::str(nm->node_names[kk ? 0 : nm->elems[k].node[0]]));
      nodes.push(J::str(nm->node_names[kk ? 0 : nm->elems[k].node[1]]));
      e.add("nodes", std::move(nodes));
      e.add("voltage", J::num(rp(k, KH_RP_VOLTAGE)));
      e.add("current", J::num(rp(k, KH_RP_CURRENT)));
      e.add("power_absorbed", J::num(rp(k, KH_RP_POWER)));
      e.add("stored_energy", J::num(rp(k, KH_RP_ENERGY)));
      e.add("model_supplied", J::num(rp(k, KH_RP_SUPPLIED)));
      els.push(std::move(e));
    }
    opdoc.add("elements", std::move(els));
    J l = J::obj();
    l.add("inflow", J::num(ledger[KH_LG_INFLOW]));
    l.add("outflow", J::num(ledger[KH_LG_OUTFLOW]));
    l.add("dissipation", J::num(ledger[KH_LG_DISSIPATION]));
    l.add("storage_rate", J::num(ledger[KH_LG_STORAGE_RATE]));
    l.add("model_supplied", J::num(ledger[KH_LG_MODEL_SUPPLIED]));
    l.add("hypothetical", J::num(ledger[KH_LG_HYPOTHETICAL]));
    l.add("residual", J::num(ledger[KH_LG_RESIDUAL]));
    l.add("relative_residual", J::num(ledger[KH_LG_RELATIVE]));
    J rows = J::arr();
    J sup = J::arr();
    for (int k = 0; k < nel; k++) {
      const int ch_ext = (int)rp(k, KH_RP_CHANNEL);
      for (auto& pt : ports_of(nm->elems[k].kind, negative(k))) {
        J r = J::obj();
        r.add("element", J::str(nm->elem_names[k]));
        r.add("channel", J::str(channel_name(pt.first)));
        r.add("status", J::str("modeled"));
        r.add("power_w", J::num(pt.first == ch_ext ? rp(k, KH_RP_EXTERIOR) : 0.0));
        r.add("note", J::str(pt.second));
        rows.push(std::move(r));
      }
      if (rp(k, KH_RP_SUPPLIED) != 0) {
        J s = J::obj();
        s.add("Element", J::str(nm->elem_names[k]));
        s.add("PowerW", J::num(rp(k, KH_RP_SUPPLIED)));
        sup.push(std::move(s));
      }
    }
    l.add("rows", std::move(rows));
    if (!sup.a.empty()) l.add("supplied", std::move(sup));
    opdoc.add("ledger", std::move(l));
  }
  J d = document(nm, version, status, have ? &opdoc : nullptr, "op");
  std::string out;
  put(out, d, 0);
  out += '\n';
  return emit(out, buf, cap);
}

int64_t kh_format_dc(int32_t format, const kh_names* nm, const char* version, kh_status status, int32_t n,
                     const kh_dc_spec* spec, const char* const* sources, int32_t points, const double* sweep_x,
                     const double* sweep_i, double max_rel, char* buf, int64_t cap) {
  if (!names_ok(nm) || (format != KH_FORMAT_JSON && format != KH_FORMAT_CSV)) return -1;
  if (!spec || spec->nsweeps < 1 || spec->nsweeps > 2 || !sources || n < nm->num_nodes - 1) return -1;
  for (int s = 0; s < spec->nsweeps; s++)
    if (!sources[s]) return -1;
  const bool have = status == KH_OK;
  if (have && (points < 0 || (points > 0 && (!sweep_x || (nm->n_elems > 0 && !sweep_i))))) return -1;
  const int nel = nm->n_elems, ns = spec->nsweeps;
  // DCSweep.values (analysis/dc.go:31-46): point k is inner index k % count0, outer k / count0
  const double cnt0 = (spec->stop[0] - spec->start[0]) / spec->step[0];
  const int64_t count0 = (int64_t)std::floor(cnt0 + 1e-9) + 1;
  auto value = [&](int s, int k) {
    const int64_t idx = s == 0 ? k % count0 : k / count0;
    return spec->start[s] + (double)idx * spec->step[s];
  };
  auto xv = [&](int k, int node) { return sweep_x[(int64_t)k * n + node - 1]; };
  auto iv = [&](int k, int e) { return sweep_i[(int64_t)k * nel + e]; };
  if (format == KH_FORMAT_CSV) {  // output.WriteDCCSV (output/ac.go:125-165)
    if (!have) return emit("", buf, cap);
    std::string out;
    std::vector<std::string> head;
    for (int s = 0; s < ns; s++) head.push_back(lower(sources[s]));
    for (int i = 1; i < nm->num_nodes; i++) head.push_back(std::string("v(") + nm->node_names[i] + ")");
    for (int e = 0; e < nel; e++) head.push_back("i(" + lower(nm->elem_names[e]) + ")");
    auto row = [&](const std::vector<std::string>& r) {
      for (size_t k = 0; k < r.size(); k++) {
        if (k) out += ',';
        csv_field(out, r[k]);
      }
      out += '\n';
    };
    row(head);
    for (int k = 0; k < points; k++) {
      std::vector<std::string> r;
      for (int s = 0; s < ns; s++) r.push_back(go_g(value(s, k), 10));
      for (int i = 1; i < nm->num_nodes; i++) r.push_back(go_g(xv(k, i), 10));
      for (int e = 0; e < nel; e++) r.push_back(go_g(iv(k, e), 10));
      row(r);
    }
    return emit(out, buf, cap);
  }
  J dcdoc;
  if (have) {  // DCDocument (output/json.go:202-221)
    dcdoc = J::obj();
    J src = J::arr();
    for (int s = 0; s < ns; s++) src.push(J::str(sources[s]));
    dcdoc.add("sources", std::move(src));
    J vals = J::arr();
    for (int s = 0; s < ns; s++) {
      J a = points > 0 ? J::arr() : J::null();
      for (int k = 0; k < points; k++) a.push(J::num(value(s, k)));
      vals.push(std::move(a));
    }
    dcdoc.add("values", std::move(vals));
    auto series = [&](auto f) {
      J a = points > 0 ? J::arr() : J::null();
      for (int k = 0; k < points; k++) a.push(J::num(f(k)));
      return a;
    };
    dcdoc.add("v", node_map(nm, [&](int i) { return series([&](int k) { return xv(k, i); }); }));
    std::vector<std::pair<std::string, int>> keys;
    for (int e = 0; e < nel; e++) keys.emplace_back(lower(nm->elem_names[e]), e);
    std::stable_sort(keys.begin(), keys.end(), [](const auto& a, const auto& b) {
      const int c = std::memcmp(a.first.data(), b.first.data(), std::min(a.first.size(), b.first.size()));
      return c != 0 ? c < 0 : a.first.size() < b.first.size();
    });
    J im = J::obj();
    for (size_t q = 0; q < keys.size(); q++) {
      if (q + 1 < keys.size() && keys[q + 1].first == keys[q].first) continue;  // a Go map keeps the last
      const int e = keys[q].second;
      im.add(keys[q].first, series([&](int k) { return iv(k, e); }));
    }
    dcdoc.add("i", std::move(im));
    dcdoc.add("max_relative_residual", J::num(max_rel));
  }
  J d = document(nm, version, status, have ? &dcdoc : nullptr, "dc");
  std::string out;
  put(out, d, 0);
  out += '\n';
  return emit(out, buf, cap);
}

int64_t kh_format_ac(int32_t format, const kh_names* nm, const char* version, kh_status status, int32_t n, int32_t nf,
                     const double* freqs, const double* xac, const double* iac, const double* ledger_f,
                     double max_rel, char* buf, int64_t cap) {
  if (!names_ok(nm) || (format != KH_FORMAT_JSON && format != KH_FORMAT_CSV)) return -1;
  if (n < nm->num_nodes - 1 || nf < 0 || (nf > 0 && !freqs)) return -1;
  const bool have = status == KH_OK;
  if (have && nf > 0 && (!xac || !ledger_f || (nm->n_elems > 0 && !iac))) return -1;
  const int nel = nm->n_elems;
  // cmplx.Abs (math.Hypot) and PhaseDeg (cmplx.Phase·180/π, output/ac.go:66)
  auto mag = [](double re, double im) -> double {
    double p = std::fabs(re), q = std::fabs(im);
    if (std::isinf(p) || std::isinf(q)) return HUGE_VAL;
    if (std::isnan(p) || std::isnan(q)) return NAN;
    if (p < q) std::swap(p, q);
    if (p == 0) return 0.0;
    q = q / p;
    return p * std::sqrt(1 + q * q);
  };
  auto phase = [](double re, double im) { return std::atan2(im, re) * 180 / 3.141592653589793; };
  auto vre = [&](int k, int node) { return xac[((int64_t)k * n + node - 1) * 2]; };
  auto vim = [&](int k, int node) { return xac[((int64_t)k * n + node - 1) * 2 + 1]; };
  auto ire = [&](int k, int e) { return iac[((int64_t)k * nel + e) * 2]; };
  auto iim = [&](int k, int e) { return iac[((int64_t)k * nel + e) * 2 + 1]; };
  auto led = [&](int k, int q) { return ledger_f[(int64_t)k * 6 + q]; };
  if (format == KH_FORMAT_CSV) {  // output.WriteACCSV (output/ac.go:67-100)
    if (!have) return emit("", buf, cap);
    std::string out;
    auto row = [&](const std::vector<std::string>& r) {
      for (size_t k = 0; k < r.size(); k++) {
        if (k) out += ',';
        csv_field(out, r[k]);
      }
      out += '\n';
    };
    std::vector<std::string> head = {"freq"};
    for (int i = 1; i < nm->num_nodes; i++) {
      head.push_back(std::string("vm(") + nm->node_names[i] + ")");
      head.push_back(std::string("vp(") + nm->node_names[i] + ")");
    }
    for (const char* h : {"ledger.inflow", "ledger.dissipation", "ledger.reactive"}) head.push_back(h);
    row(head);
    for (int k = 0; k < nf; k++) {
      std::vector<std::string> r = {go_g(freqs[k], 10)};
      for (int i = 1; i < nm->num_nodes; i++) {
        r.push_back(go_g(mag(vre(k, i), vim(k, i)), 10));
        r.push_back(go_g(phase(vre(k, i), vim(k, i)), 10));
      }
      r.push_back(go_g(led(k, 0), 10));
      r.push_back(go_g(led(k, 2), 10));
      r.push_back(go_g(led(k, 3), 10));
      row(r);
    }
    return emit(out, buf, cap);
  }
  J acdoc;
  if (have) {  // ACDocument (output/json.go:223-253)
    acdoc = J::obj();
    J fr = nf > 0 ? J::arr() : J::null();
    for (int k = 0; k < nf; k++) fr.push(J::num(freqs[k]));
    acdoc.add("freq", std::move(fr));
    auto phasors = [&](auto re, auto im) {
      J p = J::obj(), m = J::arr(), ph = J::arr();
      for (int k = 0; k < nf; k++) {
        m.push(J::num(mag(re(k), im(k))));
        ph.push(J::num(phase(re(k), im(k))));
      }
      p.add("mag", std::move(m));
      p.add("phase_deg", std::move(ph));
      return p;
    };
    acdoc.add("v", node_map(nm, [&](int i) {
      return phasors([&](int k) { return vre(k, i); }, [&](int k) { return vim(k, i); });
    }));
    std::vector<std::pair<std::string, int>> keys;
    for (int e = 0; e < nel; e++) keys.emplace_back(lower(nm->elem_names[e]), e);
    std::stable_sort(keys.begin(), keys.end(), [](const auto& a, const auto& b) {
      const int c = std::memcmp(a.first.data(), b.first.data(), std::min(a.first.size(), b.first.size()));
      return c != 0 ? c < 0 : a.first.size() < b.first.size();
    });
    J im = J::obj();
    for (size_t q = 0; q < keys.size(); q++) {
      if (q + 1 < keys.size() && keys[q + 1].first == keys[q].first) continue;  // a Go map keeps the last
      const int e = keys[q].second;
      im.add(keys[q].first, phasors([&](int k) { return ire(k, e); }, [&](int k) { return iim(k, e); }));
    }
    acdoc.add("i", std::move(im));
    J l = J::obj();
    const char* names[6] = {"inflow", "outflow", "dissipation", "reactive", "model_supplied", "residual"};
    for (int q = 0; q < 6; q++) {
      J a = J::arr();
      for (int k = 0; k < nf; k++) a.push(J::num(led(k, q)));
      l.add(names[q], std::move(a));
    }
    l.add("max_relative_residual", J::num(max_rel));
    acdoc.add("ledger", std::move(l));
  }
  J d = document(nm, version, status, have ? &acdoc : nullptr, "ac");
  std::string out;
  put(out, d, 0);
  out += '\n';
  return emit(out, buf, cap);
}

}  // extern "C"
