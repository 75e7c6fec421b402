#include "model_io.hpp"

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <stdexcept>

namespace lk {
namespace io {

const Json* Json::find(const std::string& key) const {
  for (const auto& kv : obj)
    if (kv.first == key) return &kv.second;
  return nullptr;
}

const Json& Json::at(const std::string& key) const {
  const Json* v = find(key);
  if (v == nullptr) throw std::runtime_error("model file: missing field '" + key + "'");
  return *v;
}

double Json::as_number() const {
  if (type != Number) throw std::runtime_error("model file: a number was expected");
  return num;
}

bool Json::as_bool() const {
  if (type != Bool) throw std::runtime_error("model file: a boolean was expected");
  return b;
}

const std::string& Json::as_string() const {
  if (type != String) throw std::runtime_error("model file: a string was expected");
  return str;
}

namespace {

struct Parser {
  const std::string& s;
  size_t i = 0;

  [[noreturn]] void fail(const char* what) const {
    throw std::runtime_error(std::string("model file: malformed JSON (") + what + " at offset " + std::to_string(i) +
                             ")");
  }
  void ws() {
    while (i < s.size() && (s[i] == ' ' || s[i] == '\n' || s[i] == '\r' || s[i] == '\t')) ++i;
  }
  bool eat(char c) {
    ws();
    if (i < s.size() && s[i] == c) {
      ++i;
      return true;
    }
    return false;
  }
  void expect(char c) {
    if (!eat(c)) fail("unexpected character");
  }
  std::string string_lit() {
    ws();
    if (i >= s.size() || s[i] != '"') fail("string expected");
    ++i;
    std::string out;
    while (i < s.size() && s[i] != '"') {
      char c = s[i++];
      if (c != '\\') {
        out += c;
        continue;
      }
      if (i >= s.size()) fail("bad escape");
      const char e = s[i++];
      switch (e) {
        case '"': out += '"'; break;
        case '\\': out += '\\'; break;
        case '/': out += '/'; break;
        case 'b': out += '\b'; break;
        case 'f': out += '\f'; break;
        case 'n': out += '\n'; break;
        case 'r': out += '\r'; break;
        case 't': out += '\t'; break;
        case 'u': {
          if (i + 4 > s.size()) fail("bad unicode escape");
          const unsigned cp = (unsigned)std::stoul(s.substr(i, 4), nullptr, 16);
          i += 4;
          if (cp < 0x80) {
            out += (char)cp;
          } else if (cp < 0x800) {
            out += (char)(0xC0 | (cp >> 6));
            out += (char)(0x80 | (cp & 0x3F));
          } else {
            out += (char)(0xE0 | (cp >> 12));
            out += (char)(0x80 | ((cp >> 6) & 0x3F));
            out += (char)(0x80 | (cp & 0x3F));
          }
          break;
        }
        default: fail("bad escape");
      }
    }
    if (i >= s.size()) fail("unterminated string");
    ++i;
    return out;
  }
  Json value() {
    ws();
    if (i >= s.size()) fail("value expected");
    Json v;
    const char c = s[i];
    if (c == '{') {
      ++i;
      v.type = Json::Object;
      if (eat('}')) return v;
      do {
        std::string k = string_lit();
        expect(':');
        v.obj.emplace_back(std::move(k), value());
      } while (eat(','));
      expect('}');
    } else if (c == '[') {
      ++i;
      v.type = Json::Array;
      if (eat(']')) return v;
      do {
        v.arr.push_back(value());
      } while (eat(','));
      expect(']');
    } else if (c == '"') {
      v.type = Json::String;
      v.str = string_lit();
    } else if (s.compare(i, 4, "true") == 0) {
      i += 4;
      v.type = Json::Bool;
      v.b = true;
    } else if (s.compare(i, 5, "false") == 0) {
      i += 5;
      v.type = Json::Bool;
    } else if (s.compare(i, 4, "null") == 0) {
      i += 4;
    } else {
      const char* start = s.c_str() + i;
      char* end = nullptr;
      v.num = std::strtod(start, &end);
      if (end == start) fail("number expected");
      v.type = Json::Number;
      i += (size_t)(end - start);
    }
    return v;
  }
};

const char kB64[] = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";

}  // namespace

Json parse(const std::string& text) {
  Parser p{text};
  Json v = p.value();
  p.ws();
  if (p.i != text.size()) p.fail("trailing characters");
  return v;
}

std::string base64_encode(const unsigned char* d, size_t n) {
  std::string out;
  out.reserve((n + 2) / 3 * 4);
  size_t k = 0;
  for (; k + 3 <= n; k += 3) {
    const uint32_t v = ((uint32_t)d[k] << 16) | ((uint32_t)d[k + 1] << 8) | d[k + 2];
    out += kB64[(v >> 18) & 63];
    out += kB64[(v >> 12) & 63];
    out += kB64[(v >> 6) & 63];
    out += kB64[v & 63];
  }
  if (k < n) {
    uint32_t v = (uint32_t)d[k] << 16;
    if (k + 1 < n) v |= (uint32_t)d[k + 1] << 8;
    out += kB64[(v >> 18) & 63];
    out += kB64[(v >> 12) & 63];
    out += (k + 1 < n) ? kB64[(v >> 6) & 63] : '=';
    out += '=';
  }
  return out;
}

std::vector<unsigned char> base64_decode(const std::string& s) {
  int inv[256];
  for (int& x : inv) x = -1;
  for (int c = 0; c < 64; ++c) inv[(unsigned char)kB64[c]] = c;
  std::vector<unsigned char> out;
  out.reserve(s.size() / 4 * 3);
  uint32_t acc = 0;
  int bits = 0;
  for (char ch : s) {
    if (ch == '=' || ch == '\n' || ch == '\r' || ch == ' ') continue;
    const int v = inv[(unsigned char)ch];
    if (v < 0) throw std::runtime_error("model file: bad base64 data");
    acc = (acc << 6) | (uint32_t)v;
    bits += 6;
    if (bits >= 8) {
      bits -= 8;
      out.push_back((unsigned char)((acc >> bits) & 0xFF));
    }
  }
  return out;
}

namespace {
// little-endian bytes of the doubles (the format's byte order)
std::string encode_doubles(const double* data, size_t n) {
  std::vector<unsigned char> bytes(n * sizeof(double));
  for (size_t k = 0; k < n; ++k) {
    uint64_t u;
    std::memcpy(&u, data + k, sizeof(u));
    for (int b = 0; b < 8; ++b) bytes[k * 8 + b] = (unsigned char)((u >> (8 * b)) & 0xFF);
  }
  return base64_encode(bytes.data(), bytes.size());
}
}  // namespace

std::string mat_node(const double* data, size_t n_rows, size_t n_cols) {
  return "{\"type\": \"mat\", \"n_rows\": " + std::to_string(n_rows) + ", \"n_cols\": " + std::to_string(n_cols) +
         ", \"base64_data\": \"" + encode_doubles(data, n_rows * n_cols) + "\"}";
}

std::string vec_node(const char* type, const double* data, size_t n) {
  return std::string("{\"type\": \"") + type + "\", \"size\": " + std::to_string(n) + ", \"base64_data\": \"" +
         encode_doubles(data, n) + "\"}";
}

std::vector<double> array_from(const Json& node, size_t* rows, size_t* cols) {
  const std::string& type = node.at("type").as_string();
  const std::vector<unsigned char> bytes = base64_decode(node.at("base64_data").as_string());
  if (bytes.size() % 8 != 0) throw std::runtime_error("model file: array data is not a whole number of doubles");
  const size_t n = bytes.size() / 8;
  std::vector<double> out(n);
  for (size_t k = 0; k < n; ++k) {
    uint64_t u = 0;
    for (int b = 0; b < 8; ++b) u |= (uint64_t)bytes[k * 8 + b] << (8 * b);
    std::memcpy(&out[k], &u, sizeof(double));
  }
  size_t r, c;
  if (type == "mat") {
    r = (size_t)node.at("n_rows").as_number();
    c = (size_t)node.at("n_cols").as_number();
  } else if (type == "rowvec") {
    r = 1;
    c = (size_t)node.at("size").as_number();
  } else if (type == "colvec") {
    r = (size_t)node.at("size").as_number();
    c = 1;
  } else {
    throw std::runtime_error("model file: unknown array type '" + type + "'");
  }
  if (r * c != n) throw std::runtime_error("model file: array size does not match its data");
  if (rows) *rows = r;
  if (cols) *cols = c;
  return out;
}

std::string quote(const std::string& s) {
  std::string out = "\"";
  for (char c : s) {
    if (c == '"' || c == '\\') {
      out += '\\';
      out += c;
    } else if ((unsigned char)c < 0x20) {
      char buf[8];
      std::snprintf(buf, sizeof(buf), "\\u%04x", (unsigned)(unsigned char)c);
      out += buf;
    } else {
      out += c;
    }
  }
  return out + "\"";
}

std::string number(double v) {
  if (!std::isfinite(v)) return "null";
  char buf[32];
  std::snprintf(buf, sizeof(buf), "%.17g", v);
  return buf;
}

}  // namespace io
}  // namespace lk
