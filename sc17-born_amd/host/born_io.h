// Hypercube files and program parameters for the command-line programs
// (sc17-born_amd/cli): the reference reads and writes its volumes through
// SEP genericIO (ioModes/genericRegFile, e.g. rtm.cc:11-22, util/Add.cc:5-40)
// and takes its parameters as key=value arguments or a parameter file
// (scripts/run.py:41-52 writes `json=mig.P`).
//
// Two descriptor formats, chosen by the descriptor's file name:
//  - *.json (genericIO JSON): one object with n1, o1, d1, label1, ... and
//    "filename", the raw data file (scripts/run.py:18-35); the data are
//    native (little-endian) float32.  Written with filename = descriptor +
//    ".dat".
//  - anything else (SEP, *.H): text of key=value pairs, the last of each key
//    winning; in= names the data file, data_format= "xdr_float" (big-endian,
//    the SEP default) or "native_float".  Written as native_float with
//    in = descriptor + "@".
// Data are x fastest: axis 1 varies fastest.
#pragma once

#include <map>
#include <string>
#include <vector>

namespace born_io {

struct Axis {
  int n = 1;
  float o = 0.f, d = 1.f;
  std::string label;
};

struct Cube {
  std::vector<Axis> axes;
  std::vector<float> vals;
  long long n123() const;
};

// Parameters from key=value arguments; json=FILE adds the top-level keys of
// a JSON object, par=FILE the key=value pairs of a text file.  Arguments
// given later override earlier ones.  Getters throw std::runtime_error when
// a key without a default is missing or a value does not parse.
class Params {
 public:
  Params(int argc, char **argv);
  bool has(const std::string &k) const { return kv_.count(k) != 0; }
  std::string str(const std::string &k) const;
  std::string str(const std::string &k, const std::string &def) const;
  int i(const std::string &k, int def) const;
  float f(const std::string &k, float def) const;
  void set(const std::string &k, const std::string &v) { kv_[k] = v; }

 private:
  std::map<std::string, std::string> kv_;
};

// Flat key/value view of a JSON object's top-level scalars (numbers and
// booleans as written, strings unquoted); nested values are skipped.
std::map<std::string, std::string> parse_json_object(const std::string &text);
// key=value pairs of a SEP header or parameter text ('#' starts a comment)
std::map<std::string, std::string> parse_kv_text(const std::string &text);

// axes and data of a descriptor (the data file resolved next to it)
Cube read_cube(const std::string &desc);
// axes only
std::vector<Axis> read_axes(const std::string &desc);
// write descriptor and data
void write_cube(const std::string &desc, const Cube &c);

}  // namespace born_io
