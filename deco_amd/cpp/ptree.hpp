// Minimal property tree with the subset of boost::property_tree::ptree that deco uses:
// get<T>("a.b"), get_child("a"), put, and a JSON reader that strips // and /* */ comments first
// (IO/ptree.hpp:27-119). Leaves are strings converted on access, as in Boost.
#pragma once
#include <cctype>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace deco {

class ptree {
 public:
  std::string data;
  std::vector<std::pair<std::string, ptree>> children;

  const ptree* find(const std::string& key) const {
    for (auto& kv : children)
      if (kv.first == key) return &kv.second;
    return nullptr;
  }
  ptree* find(const std::string& key) {
    for (auto& kv : children)
      if (kv.first == key) return &kv.second;
    return nullptr;
  }
  const ptree& get_child(const std::string& path) const {
    const ptree* node = this;
    std::string::size_type start = 0;
    while (start <= path.size()) {
      auto dot = path.find('.', start);
      std::string key = path.substr(start, dot == std::string::npos ? std::string::npos : dot - start);
      node = node->find(key);
      if (!node) throw std::runtime_error("No such node (" + path + ")");
      if (dot == std::string::npos) break;
      start = dot + 1;
    }
    return *node;
  }
  template <class T>
  T get(const std::string& path) const {
    const std::string& s = get_child(path).data;
    std::istringstream is(s);
    T v{};
    is >> v;
    if (is.fail()) throw std::runtime_error("conversion of data to type failed (" + path + " = \"" + s + "\")");
    return v;
  }
  template <class T>
  T get(const std::string& path, const T& fallback) const {
    try { return get<T>(path); } catch (const std::exception&) { return fallback; }
  }
  void put(const std::string& path, const std::string& value) {
    ptree* node = this;
    std::string::size_type start = 0;
    while (true) {
      auto dot = path.find('.', start);
      std::string key = path.substr(start, dot == std::string::npos ? std::string::npos : dot - start);
      ptree* next = node->find(key);
      if (!next) {
        node->children.emplace_back(key, ptree());
        next = &node->children.back().second;
      }
      node = next;
      if (dot == std::string::npos) break;
      start = dot + 1;
    }
    node->data = value;
  }
};

template <>
inline std::string ptree::get<std::string>(const std::string& path) const {
  return get_child(path).data;
}

namespace json_detail {
struct Parser {
  const std::string& s;
  size_t i = 0;
  explicit Parser(const std::string& t) : s(t) {}
  void ws() { while (i < s.size() && std::isspace((unsigned char)s[i])) ++i; }
  [[noreturn]] void err(const char* m) { throw std::runtime_error(std::string("JSON parse error: ") + m); }
  std::string str() {
    if (s[i] != '"') err("expected string");
    ++i;
    std::string out;
    while (i < s.size() && s[i] != '"') {
      if (s[i] == '\\' && i + 1 < s.size()) { out += s[i + 1]; i += 2; }
      else out += s[i++];
    }
    if (i >= s.size()) err("unterminated string");
    ++i;
    return out;
  }
  void value(ptree& node) {
    ws();
    if (i >= s.size()) err("unexpected end");
    if (s[i] == '{') {
      ++i;
      ws();
      if (s[i] == '}') { ++i; return; }
      while (true) {
        ws();
        std::string k = str();
        ws();
        if (s[i] != ':') err("expected ':'");
        ++i;
        node.children.emplace_back(k, ptree());
        value(node.children.back().second);
        ws();
        if (s[i] == ',') { ++i; continue; }
        if (s[i] == '}') { ++i; return; }
        err("expected ',' or '}'");
      }
    } else if (s[i] == '[') {
      ++i;
      ws();
      if (s[i] == ']') { ++i; return; }
      while (true) {
        node.children.emplace_back("", ptree());
        value(node.children.back().second);
        ws();
        if (s[i] == ',') { ++i; continue; }
        if (s[i] == ']') { ++i; return; }
        err("expected ',' or ']'");
      }
    } else if (s[i] == '"') {
      node.data = str();
    } else {
      size_t j = i;
      while (j < s.size() && s[j] != ',' && s[j] != '}' && s[j] != ']' && !std::isspace((unsigned char)s[j])) ++j;
      node.data = s.substr(i, j - i);
      i = j;
    }
  }
};
}  // namespace json_detail

// IO::remove_json_comments (IO/ptree.hpp:27-84): per line, no string-literal tracking
inline std::string remove_json_comments(const std::string& input) {
  std::istringstream in(input);
  std::ostringstream out;
  std::string line;
  bool in_block = false;
  while (std::getline(in, line)) {
    std::string piece;
    size_t pos = 0;
    while (pos < line.size()) {
      if (in_block) {
        auto e = line.find("*/", pos);
        if (e == std::string::npos) { pos = line.size(); break; }
        in_block = false;
        pos = e + 2;
      }
      auto sl = line.find("//", pos), bl = line.find("/*", pos);
      if (sl != std::string::npos && (bl == std::string::npos || sl < bl)) { piece += line.substr(pos, sl - pos); break; }
      if (bl != std::string::npos) { piece += line.substr(pos, bl - pos); in_block = true; pos = bl + 2; continue; }
      piece += line.substr(pos);
      break;
    }
    out << piece << "\n";
  }
  return out.str();
}

inline ptree read_json_string(const std::string& text) {
  ptree pt;
  std::string clean = remove_json_comments(text);
  json_detail::Parser p(clean);
  p.value(pt);
  return pt;
}

// IO::read_ptree(name) (IO/ptree.hpp:121-149)
inline ptree read_ptree(const char* name) {
  std::ifstream f(name);
  if (!f) throw std::runtime_error(std::string("Could not open file ") + name);
  std::string text((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
  return read_json_string(text);
}

}  // namespace deco
