// Model files: the reference's JSON "version 2" Kriging format
// (Kriging::save / Kriging::load, src/lib/Kriging.cpp:2691-2750,
// KrigingImpl::dump_common_to_json / load_common_from_json,
// src/lib/KrigingImpl.cpp:644-711): one JSON object whose matrices are
// {"type": "mat" | "rowvec" | "colvec", sizes, "base64_data"}, the data being
// the little-endian IEEE doubles of the column-major array in standard base64
// (src/lib/utils/jsonutils.cpp:66-110).  A small JSON reader / writer of our
// own; only what these files use (objects, arrays, strings, numbers, bools,
// null).
#pragma once
#include <string>
#include <utility>
#include <vector>

namespace lk {
namespace io {

struct Json {
  enum Type { Null, Bool, Number, String, Array, Object };
  Type type = Null;
  bool b = false;
  double num = 0.0;
  std::string str;
  std::vector<Json> arr;
  std::vector<std::pair<std::string, Json>> obj;

  const Json* find(const std::string& key) const;
  const Json& at(const std::string& key) const;  // throws when missing
  double as_number() const;
  bool as_bool() const;
  const std::string& as_string() const;
};

Json parse(const std::string& text);

std::string base64_encode(const unsigned char* data, size_t n);
std::vector<unsigned char> base64_decode(const std::string& s);

// encoded array node (value text) and its decoding; n_cols = 1 for vectors
std::string mat_node(const double* data, size_t n_rows, size_t n_cols);
std::string vec_node(const char* type, const double* data, size_t n);
// doubles of a "mat" / "rowvec" / "colvec" node; sets rows / cols
std::vector<double> array_from(const Json& node, size_t* rows, size_t* cols);

std::string quote(const std::string& s);
std::string number(double v);

}  // namespace io
}  // namespace lk
