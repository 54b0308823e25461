// json_min.h -- a small JSON reader for the run_config files.
//
// The reference reads its run_config with nlohmann::json
// (DataKeeper/src/datakeeper.cpp:6-48); only objects, arrays, strings,
// numbers, true/false/null are needed here.  Parse errors throw
// std::runtime_error with the byte offset.
#pragma once

#include <cctype>
#include <cstdlib>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace femproject {

struct Json {
  enum Kind { NUL, BOOL, NUM, STR, ARR, OBJ } kind = NUL;
  bool b = false;
  double num = 0.0;
  std::string str;
  std::vector<Json> arr;
  std::map<std::string, Json> obj;

  bool has(const std::string &k) const { return kind == OBJ && obj.count(k) != 0; }
  const Json &at(const std::string &k) const {
    auto it = obj.find(k);
    if (kind != OBJ || it == obj.end()) throw std::runtime_error("missing key: " + k);
    return it->second;
  }
  double number_or(const std::string &k, double d) const {
    return has(k) && at(k).kind == NUM ? at(k).num : d;
  }
  std::string string_or(const std::string &k, const std::string &d) const {
    return has(k) && at(k).kind == STR ? at(k).str : d;
  }
  bool bool_or(const std::string &k, bool d) const {
    return has(k) && at(k).kind == BOOL ? at(k).b : d;
  }

  static Json parse(const std::string &s) {
    size_t i = 0;
    Json v = value(s, i);
    ws(s, i);
    if (i != s.size()) fail(i);
    return v;
  }

 private:
  [[noreturn]] static void fail(size_t i) {
    throw std::runtime_error("JSON parse error at byte " + std::to_string(i));
  }
  static void ws(const std::string &s, size_t &i) {
    while (i < s.size() && std::isspace(static_cast<unsigned char>(s[i]))) ++i;
  }
  static Json value(const std::string &s, size_t &i) {
    ws(s, i);
    if (i >= s.size()) fail(i);
    Json v;
    const char c = s[i];
    if (c == '{') {
      v.kind = OBJ;
      ++i;
      ws(s, i);
      if (i < s.size() && s[i] == '}') {
        ++i;
        return v;
      }
      for (;;) {
        ws(s, i);
        if (i >= s.size() || s[i] != '"') fail(i);
        std::string k = string(s, i);
        ws(s, i);
        if (i >= s.size() || s[i] != ':') fail(i);
        ++i;
        v.obj[k] = value(s, i);
        ws(s, i);
        if (i < s.size() && s[i] == ',') {
          ++i;
          continue;
        }
        if (i < s.size() && s[i] == '}') {
          ++i;
          return v;
        }
        fail(i);
      }
    }
    if (c == '[') {
      v.kind = ARR;
      ++i;
      ws(s, i);
      if (i < s.size() && s[i] == ']') {
        ++i;
        return v;
      }
      for (;;) {
        v.arr.push_back(value(s, i));
        ws(s, i);
        if (i < s.size() && s[i] == ',') {
          ++i;
          continue;
        }
        if (i < s.size() && s[i] == ']') {
          ++i;
          return v;
        }
        fail(i);
      }
    }
    if (c == '"') {
      v.kind = STR;
      v.str = string(s, i);
      return v;
    }
    if (s.compare(i, 4, "true") == 0) {
      v.kind = BOOL;
      v.b = true;
      i += 4;
      return v;
    }
    if (s.compare(i, 5, "false") == 0) {
      v.kind = BOOL;
      i += 5;
      return v;
    }
    if (s.compare(i, 4, "null") == 0) {
      i += 4;
      return v;
    }
    const char *p = s.c_str() + i;
    char *end = nullptr;
    v.num = std::strtod(p, &end);
    if (end == p) fail(i);
    v.kind = NUM;
    i += size_t(end - p);
    return v;
  }
  static std::string string(const std::string &s, size_t &i) {
    ++i;  // opening quote
    std::string out;
    while (i < s.size() && s[i] != '"') {
      if (s[i] == '\\') {
        if (++i >= s.size()) fail(i);
        const char e = s[i];
        if (e == 'n')
          out += '\n';
        else if (e == 't')
          out += '\t';
        else if (e == 'u') {
          if (i + 4 >= s.size()) fail(i);
          out += '?';
          i += 4;
        } else
          out += e;
        ++i;
      } else {
        out += s[i++];
      }
    }
    if (i >= s.size()) fail(i);
    ++i;  // closing quote
    return out;
  }
};

}  // namespace femproject
