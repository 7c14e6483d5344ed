// Hypercube files and program parameters (see born_io.h).
#include "born_io.h"

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace born_io {

long long Cube::n123() const {
  long long n = 1;
  for (const Axis &a : axes) n *= a.n;
  return n;
}

static std::string slurp(const std::string &path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error("cannot open " + path);
  std::ostringstream s;
  s << f.rdbuf();
  return s.str();
}

static bool ends_with(const std::string &s, const std::string &t) {
  return s.size() >= t.size() && s.compare(s.size() - t.size(), t.size(), t) == 0;
}

static std::string dir_of(const std::string &path) {
  const size_t k = path.find_last_of('/');
  return k == std::string::npos ? std::string() : path.substr(0, k + 1);
}

static std::string base_of(const std::string &path) {
  const size_t k = path.find_last_of('/');
  return k == std::string::npos ? path : path.substr(k + 1);
}

static bool exists(const std::string &path) {
  std::ifstream f(path);
  return (bool)f;
}

// ------------------------------------------------------------------- JSON

namespace {
struct JsonReader {
  const std::string &t;
  size_t p = 0;
  explicit JsonReader(const std::string &text) : t(text) {}
  [[noreturn]] void fail(const char *what) const {
    throw std::runtime_error(std::string("JSON: ") + what + " at offset " + std::to_string(p));
  }
  void ws() {
    while (p < t.size() && (t[p] == ' ' || t[p] == '\t' || t[p] == '\n' || t[p] == '\r')) p++;
  }
  std::string string() {
    if (t[p] != '"') fail("expected a string");
    p++;
    std::string out;
    while (p < t.size() && t[p] != '"') {
      char c = t[p++];
      if (c == '\\') {
        if (p >= t.size()) fail("bad escape");
        const char e = t[p++];
        switch (e) {
          case 'n': c = '\n'; break;
          case 't': c = '\t'; break;
          case 'r': c = '\r'; break;
          case 'b': c = '\b'; break;
          case 'f': c = '\f'; break;
          case 'u':  // kept as written: descriptors hold ASCII names
            out += "\\u";
            continue;
          default: c = e;
        }
      }
      out += c;
    }
    if (p >= t.size()) fail("unterminated string");
    p++;
    return out;
  }
  // a scalar as text, or "" after skipping an object or array
  std::string value(bool *scalar) {
    ws();
    if (p >= t.size()) fail("unexpected end");
    *scalar = true;
    if (t[p] == '"') return string();
    if (t[p] == '{' || t[p] == '[') {
      *scalar = false;
      skip_nested();
      return std::string();
    }
    const size_t b = p;
    while (p < t.size() && t[p] != ',' && t[p] != '}' && t[p] != ']' && t[p] != ' ' &&
           t[p] != '\n' && t[p] != '\r' && t[p] != '\t')
      p++;
    if (p == b) fail("expected a value");
    return t.substr(b, p - b);
  }
  void skip_nested() {
    int depth = 0;
    do {
      if (p >= t.size()) fail("unterminated object or array");
      if (t[p] == '"') {
        string();
        continue;
      }
      if (t[p] == '{' || t[p] == '[') depth++;
      if (t[p] == '}' || t[p] == ']') depth--;
      p++;
    } while (depth > 0);
  }
};
}  // namespace

std::map<std::string, std::string> parse_json_object(const std::string &text) {
  JsonReader r(text);
  std::map<std::string, std::string> kv;
  r.ws();
  if (r.p >= text.size() || text[r.p] != '{') r.fail("expected an object");
  r.p++;
  r.ws();
  if (r.p < text.size() && text[r.p] == '}') return kv;
  while (true) {
    r.ws();
    const std::string key = r.string();
    r.ws();
    if (r.p >= text.size() || text[r.p] != ':') r.fail("expected ':'");
    r.p++;
    bool scalar = false;
    const std::string v = r.value(&scalar);
    if (scalar) kv[key] = v;
    r.ws();
    if (r.p >= text.size()) r.fail("unterminated object");
    if (text[r.p] == ',') {
      r.p++;
      continue;
    }
    if (text[r.p] == '}') break;
    r.fail("expected ',' or '}'");
  }
  return kv;
}

// ------------------------------------------------------------ key=value text

std::map<std::string, std::string> parse_kv_text(const std::string &text) {
  std::map<std::string, std::string> kv;
  size_t p = 0;
  const size_t n = text.size();
  auto space = [](char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r'; };
  while (p < n) {
    while (p < n && space(text[p])) p++;
    if (p >= n) break;
    if (text[p] == '#') {  // comment to the end of the line
      while (p < n && text[p] != '\n') p++;
      continue;
    }
    const size_t b = p;
    while (p < n && !space(text[p]) && text[p] != '=') p++;
    if (p >= n || text[p] != '=') {  // a word without '=': skipped
      while (p < n && !space(text[p])) p++;
      continue;
    }
    const std::string key = text.substr(b, p - b);
    p++;
    std::string v;
    if (p < n && (text[p] == '"' || text[p] == '\'')) {
      const char q = text[p++];
      const size_t vb = p;
      while (p < n && text[p] != q) p++;
      v = text.substr(vb, p - vb);
      if (p < n) p++;
    } else {
      const size_t vb = p;
      while (p < n && !space(text[p])) p++;
      v = text.substr(vb, p - vb);
    }
    if (!key.empty()) kv[key] = v;
  }
  return kv;
}

// -------------------------------------------------------------- parameters

Params::Params(int argc, char **argv) {
  for (int a = 1; a < argc; a++) {
    const std::string arg = argv[a];
    const size_t eq = arg.find('=');
    if (eq == std::string::npos || eq == 0) continue;
    const std::string k = arg.substr(0, eq), v = arg.substr(eq + 1);
    if (k == "json") {
      for (auto &e : parse_json_object(slurp(v))) kv_[e.first] = e.second;
    } else if (k == "par") {
      for (auto &e : parse_kv_text(slurp(v))) kv_[e.first] = e.second;
    } else {
      kv_[k] = v;
    }
  }
}

std::string Params::str(const std::string &k) const {
  auto it = kv_.find(k);
  if (it == kv_.end()) throw std::runtime_error("missing parameter " + k + "=");
  return it->second;
}

std::string Params::str(const std::string &k, const std::string &def) const {
  auto it = kv_.find(k);
  return it == kv_.end() ? def : it->second;
}

int Params::i(const std::string &k, int def) const {
  auto it = kv_.find(k);
  if (it == kv_.end()) return def;
  char *e = nullptr;
  const long v = std::strtol(it->second.c_str(), &e, 10);
  if (e == it->second.c_str() || *e) throw std::runtime_error("bad integer " + k + "=" + it->second);
  return (int)v;
}

float Params::f(const std::string &k, float def) const {
  auto it = kv_.find(k);
  if (it == kv_.end()) return def;
  char *e = nullptr;
  const float v = std::strtof(it->second.c_str(), &e);
  if (e == it->second.c_str() || *e) throw std::runtime_error("bad number " + k + "=" + it->second);
  return v;
}

// -------------------------------------------------------------- hypercubes

static std::vector<Axis> axes_from(const std::map<std::string, std::string> &kv) {
  std::vector<Axis> axes;
  auto num = [&](const std::string &k, double def) {
    auto it = kv.find(k);
    if (it == kv.end()) return def;
    char *e = nullptr;
    const double v = std::strtod(it->second.c_str(), &e);
    if (e == it->second.c_str()) throw std::runtime_error("bad value " + k + "=" + it->second);
    return v;
  };
  for (int k = 1; kv.count("n" + std::to_string(k)); k++) {
    const std::string s = std::to_string(k);
    Axis a;
    a.n = (int)num("n" + s, 1);
    a.o = (float)num("o" + s, 0.);
    a.d = (float)num("d" + s, 1.);
    auto it = kv.find("label" + s);
    if (it != kv.end()) a.label = it->second;
    if (a.n < 1) throw std::runtime_error("axis " + s + " has n < 1");
    axes.push_back(a);
  }
  if (axes.empty()) throw std::runtime_error("descriptor has no n1");
  return axes;
}

struct Desc {
  std::vector<Axis> axes;
  std::string data;  // data file as resolved
  bool big_endian = false;
};

static Desc read_desc(const std::string &desc) {
  const std::string text = slurp(desc);
  const bool json = ends_with(desc, ".json");
  const auto kv = json ? parse_json_object(text) : parse_kv_text(text);
  Desc d;
  d.axes = axes_from(kv);
  auto it = kv.find(json ? "filename" : "in");
  if (it == kv.end())
    throw std::runtime_error(desc + ": no data file (" + (json ? "filename" : "in") + "=)");
  std::string f = it->second;
  if (!f.empty() && f[0] != '/' && !exists(f) && exists(dir_of(desc) + f)) f = dir_of(desc) + f;
  d.data = f;
  if (!json) {
    auto df = kv.find("data_format");
    const std::string fmt = df == kv.end() ? "xdr_float" : df->second;
    if (fmt == "xdr_float")
      d.big_endian = true;
    else if (fmt != "native_float")
      throw std::runtime_error(desc + ": data_format " + fmt + " is not a float format");
    auto es = kv.find("esize");
    if (es != kv.end() && es->second != "4") throw std::runtime_error(desc + ": esize is not 4");
  }
  return d;
}

std::vector<Axis> read_axes(const std::string &desc) { return read_desc(desc).axes; }

Cube read_cube(const std::string &desc) {
  const Desc d = read_desc(desc);
  Cube c;
  c.axes = d.axes;
  c.vals.resize((size_t)c.n123());
  FILE *f = std::fopen(d.data.c_str(), "rb");
  if (!f) throw std::runtime_error(desc + ": cannot open data file " + d.data);
  const size_t got = std::fread(c.vals.data(), sizeof(float), c.vals.size(), f);
  std::fclose(f);
  if (got != c.vals.size())
    throw std::runtime_error(desc + ": data file " + d.data + " holds " + std::to_string(got) +
                             " of " + std::to_string(c.vals.size()) + " floats");
  if (d.big_endian)
    for (float &v : c.vals) {
      uint32_t u;
      std::memcpy(&u, &v, 4);
      u = __builtin_bswap32(u);
      std::memcpy(&v, &u, 4);
    }
  return c;
}

void write_cube(const std::string &desc, const Cube &c) {
  if ((long long)c.vals.size() != c.n123())
    throw std::runtime_error(desc + ": values do not match the axes");
  const bool json = ends_with(desc, ".json");
  std::string data_path, data_name;
  if (json) {
    data_name = base_of(desc) + ".dat";
    data_path = dir_of(desc) + data_name;
  } else {
    data_path = desc + "@";
  }
  {
    FILE *f = std::fopen(data_path.c_str(), "wb");
    if (!f) throw std::runtime_error("cannot write " + data_path);
    const size_t put = std::fwrite(c.vals.data(), sizeof(float), c.vals.size(), f);
    std::fclose(f);
    if (put != c.vals.size()) throw std::runtime_error("short write to " + data_path);
  }
  std::ofstream h(desc);
  if (!h) throw std::runtime_error("cannot write " + desc);
  char buf[64];
  auto num = [&](float v) {
    std::snprintf(buf, sizeof buf, "%.9g", v);
    return std::string(buf);
  };
  if (json) {
    h << "{\n";
    for (size_t k = 0; k < c.axes.size(); k++) {
      const std::string s = std::to_string(k + 1);
      const Axis &a = c.axes[k];
      h << "    \"n" << s << "\": " << a.n << ",\n";
      h << "    \"o" << s << "\": " << num(a.o) << ",\n";
      h << "    \"d" << s << "\": " << num(a.d) << ",\n";
      h << "    \"label" << s << "\": \"" << (a.label.empty() ? "Undefined" : a.label) << "\",\n";
    }
    h << "    \"filename\": \"" << data_name << "\"\n}\n";
  } else {
    for (size_t k = 0; k < c.axes.size(); k++) {
      const std::string s = std::to_string(k + 1);
      const Axis &a = c.axes[k];
      h << "n" << s << "=" << a.n << " o" << s << "=" << num(a.o) << " d" << s << "="
        << num(a.d);
      if (!a.label.empty()) h << " label" << s << "=\"" << a.label << "\"";
      h << "\n";
    }
    h << "in=\"" << data_path << "\"\ndata_format=\"native_float\"\nesize=4\n";
  }
  if (!h) throw std::runtime_error("short write to " + desc);
}

}  // namespace born_io
