// zg_json.h -- minimal JSON value, parser and pretty writer for manifest.json.
// Numbers keep their source text so u64 seeds and f64 fields round-trip.
#pragma once

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace zg {

struct Json {
  enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
  bool b = false;
  std::string text;  // number text or string value
  std::vector<Json> items;
  std::vector<std::pair<std::string, Json>> fields;  // insertion order

  static Json null() { return Json(); }
  static Json boolean(bool v) { Json j; j.kind = Bool; j.b = v; return j; }
  static Json str(const std::string &s) { Json j; j.kind = String; j.text = s; return j; }
  static Json u64(uint64_t v) { Json j; j.kind = Number; j.text = std::to_string(v); return j; }
  static Json f64(double v) {
    Json j;
    j.kind = Number;
    char buf[64];
    snprintf(buf, sizeof buf, "%.17g", v);
    std::string s = buf;
    if (s.find_first_of(".eEn") == std::string::npos) s += ".0";
    j.text = s;
    return j;
  }
  static Json array() { Json j; j.kind = Array; return j; }
  static Json object() { Json j; j.kind = Object; return j; }

  Json &set(const std::string &key, Json v) {
    for (auto &f : fields)
      if (f.first == key) { f.second = std::move(v); return f.second; }
    fields.emplace_back(key, std::move(v));
    return fields.back().second;
  }
  void push(Json v) { items.push_back(std::move(v)); }

  const Json *get(const std::string &key) const {
    for (auto &f : fields)
      if (f.first == key) return &f.second;
    return nullptr;
  }
  bool is_null() const { return kind == Null; }
};

class JsonError : public std::runtime_error {
 public:
  using std::runtime_error::runtime_error;
};

class JsonParser {
 public:
  explicit JsonParser(const std::string &s) : s_(s) {}
  Json parse() {
    Json v = value();
    ws();
    if (i_ != s_.size()) fail("trailing characters");
    return v;
  }

 private:
  const std::string &s_;
  size_t i_ = 0;
  int depth_ = 0;

  [[noreturn]] void fail(const char *what) {
    throw JsonError(std::string("invalid JSON: ") + what + " at byte " +
                    std::to_string(i_));
  }
  void ws() {
    while (i_ < s_.size() &&
           (s_[i_] == ' ' || s_[i_] == '\n' || s_[i_] == '\r' || s_[i_] == '\t'))
      i_++;
  }
  bool lit(const char *w) {
    size_t n = 0;
    while (w[n]) n++;
    if (s_.compare(i_, n, w) == 0) { i_ += n; return true; }
    return false;
  }
  static void utf8(std::string &out, uint32_t cp) {
    if (cp < 0x80) out += (char)cp;
    else if (cp < 0x800) { out += (char)(0xC0 | (cp >> 6)); out += (char)(0x80 | (cp & 0x3F)); }
    else if (cp < 0x10000) {
      out += (char)(0xE0 | (cp >> 12)); out += (char)(0x80 | ((cp >> 6) & 0x3F));
      out += (char)(0x80 | (cp & 0x3F));
    } else {
      out += (char)(0xF0 | (cp >> 18)); out += (char)(0x80 | ((cp >> 12) & 0x3F));
      out += (char)(0x80 | ((cp >> 6) & 0x3F)); out += (char)(0x80 | (cp & 0x3F));
    }
  }
  uint32_t hex4() {
    if (i_ + 4 > s_.size()) fail("short \\u escape");
    uint32_t v = 0;
    for (int n = 0; n < 4; n++) {
      char c = s_[i_++];
      v <<= 4;
      if (c >= '0' && c <= '9') v |= c - '0';
      else if (c >= 'a' && c <= 'f') v |= c - 'a' + 10;
      else if (c >= 'A' && c <= 'F') v |= c - 'A' + 10;
      else fail("bad \\u escape");
    }
    return v;
  }
  std::string string_body() {
    if (s_[i_] != '"') fail("expected string");
    i_++;
    std::string out;
    while (true) {
      if (i_ >= s_.size()) fail("unterminated string");
      char c = s_[i_++];
      if (c == '"') break;
      if ((unsigned char)c < 0x20) fail("control character in string");
      if (c != '\\') { out += c; continue; }
      if (i_ >= s_.size()) fail("bad escape");
      char e = s_[i_++];
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
          uint32_t cp = hex4();
          if (cp >= 0xD800 && cp < 0xDC00) {
            if (!(i_ + 1 < s_.size() && s_[i_] == '\\' && s_[i_ + 1] == 'u')) fail("lone surrogate");
            i_ += 2;
            uint32_t lo = hex4();
            if (lo < 0xDC00 || lo > 0xDFFF) fail("bad surrogate pair");
            cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
          } else if (cp >= 0xDC00 && cp < 0xE000) {
            fail("lone surrogate");
          }
          utf8(out, cp);
          break;
        }
        default: fail("bad escape");
      }
    }
    return out;
  }
  Json value() {
    ws();
    if (i_ >= s_.size()) fail("unexpected end");
    if (++depth_ > 128) fail("nesting too deep");
    Json v;
    char c = s_[i_];
    if (c == '{') {
      v.kind = Json::Object;
      i_++;
      ws();
      if (i_ < s_.size() && s_[i_] == '}') { i_++; depth_--; return v; }
      while (true) {
        ws();
        std::string key = string_body();
        ws();
        if (i_ >= s_.size() || s_[i_] != ':') fail("expected ':'");
        i_++;
        Json item = value();
        if (v.get(key)) fail("duplicate field");
        v.fields.emplace_back(std::move(key), std::move(item));
        ws();
        if (i_ < s_.size() && s_[i_] == ',') { i_++; continue; }
        if (i_ < s_.size() && s_[i_] == '}') { i_++; break; }
        fail("expected ',' or '}'");
      }
    } else if (c == '[') {
      v.kind = Json::Array;
      i_++;
      ws();
      if (i_ < s_.size() && s_[i_] == ']') { i_++; depth_--; return v; }
      while (true) {
        v.items.push_back(value());
        ws();
        if (i_ < s_.size() && s_[i_] == ',') { i_++; continue; }
        if (i_ < s_.size() && s_[i_] == ']') { i_++; break; }
        fail("expected ',' or ']'");
      }
    } else if (c == '"') {
      v.kind = Json::String;
      v.text = string_body();
    } else if (lit("true")) {
      v.kind = Json::Bool; v.b = true;
    } else if (lit("false")) {
      v.kind = Json::Bool; v.b = false;
    } else if (lit("null")) {
      v.kind = Json::Null;
    } else if (c == '-' || (c >= '0' && c <= '9')) {
      size_t start = i_;
      if (s_[i_] == '-') i_++;
      while (i_ < s_.size() && ((s_[i_] >= '0' && s_[i_] <= '9') || s_[i_] == '.' ||
                                s_[i_] == 'e' || s_[i_] == 'E' || s_[i_] == '+' ||
                                s_[i_] == '-'))
        i_++;
      v.kind = Json::Number;
      v.text = s_.substr(start, i_ - start);
    } else {
      fail("unexpected character");
    }
    depth_--;
    return v;
  }
};

inline void json_escape(std::string &out, const std::string &s) {
  out += '"';
  for (unsigned char c : s) {
    switch (c) {
      case '"': out += "\\\""; break;
      case '\\': out += "\\\\"; break;
      case '\n': out += "\\n"; break;
      case '\r': out += "\\r"; break;
      case '\t': out += "\\t"; break;
      case '\b': out += "\\b"; break;
      case '\f': out += "\\f"; break;
      default:
        if (c < 0x20) {
          char buf[8];
          snprintf(buf, sizeof buf, "\\u%04x", c);
          out += buf;
        } else {
          out += (char)c;
        }
    }
  }
  out += '"';
}

// Pretty printer with two-space indentation (serde_json::to_writer_pretty style).
inline void json_write(std::string &out, const Json &v, int indent = 0) {
  auto pad = [&](int n) { out.append((size_t)n * 2, ' '); };
  switch (v.kind) {
    case Json::Null: out += "null"; break;
    case Json::Bool: out += v.b ? "true" : "false"; break;
    case Json::Number: out += v.text; break;
    case Json::String: json_escape(out, v.text); break;
    case Json::Array:
      if (v.items.empty()) { out += "[]"; break; }
      out += "[\n";
      for (size_t i = 0; i < v.items.size(); i++) {
        pad(indent + 1);
        json_write(out, v.items[i], indent + 1);
        if (i + 1 < v.items.size()) out += ',';
        out += '\n';
      }
      pad(indent);
      out += ']';
      break;
    case Json::Object:
      if (v.fields.empty()) { out += "{}"; break; }
      out += "{\n";
      for (size_t i = 0; i < v.fields.size(); i++) {
        pad(indent + 1);
        json_escape(out, v.fields[i].first);
        out += ": ";
        json_write(out, v.fields[i].second, indent + 1);
        if (i + 1 < v.fields.size()) out += ',';
        out += '\n';
      }
      pad(indent);
      out += '}';
      break;
  }
}

}  // namespace zg
