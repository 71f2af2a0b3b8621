// Reader for the YAML subset of MISA-MD configuration files (yaml_lite.hpp).
#include "yaml_lite.hpp"

#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdlib>

namespace misa {
namespace yaml {

namespace {

struct Line {
  int no;           // 1-based line number in the file
  int indent;       // leading spaces
  std::string text; // content without indentation, comment and trailing blanks
};

bool isBlank(char c) { return c == ' ' || c == '\t'; }

// Remove a comment: '#' at the start or after a blank, outside quotes.
std::string stripComment(const std::string& s) {
  char q = 0;
  for (size_t i = 0; i < s.size(); i++) {
    const char c = s[i];
    if (q) {
      if (c == q) q = 0;
      continue;
    }
    if (c == '"' || c == '\'') {
      if (i == 0 || isBlank(s[i - 1]) || s[i - 1] == '[' || s[i - 1] == ',' || s[i - 1] == ':') q = c;
      continue;
    }
    if (c == '#' && (i == 0 || isBlank(s[i - 1]))) return s.substr(0, i);
  }
  return s;
}

std::string trim(const std::string& s) {
  size_t a = 0, b = s.size();
  while (a < b && isBlank(s[a])) a++;
  while (b > a && isBlank(s[b - 1])) b--;
  return s.substr(a, b - a);
}

// Unquote a scalar token ("..." with backslash escapes, '...' with '').
bool unquote(const std::string& t, std::string& out, std::string& err) {
  if (t.size() >= 2 && t.front() == '"' && t.back() == '"') {
    out.clear();
    for (size_t i = 1; i + 1 < t.size(); i++) {
      char c = t[i];
      if (c == '\\' && i + 2 < t.size()) {
        const char e = t[++i];
        c = e == 'n' ? '\n' : e == 't' ? '\t' : e;
      }
      out += c;
    }
    return true;
  }
  if (t.size() >= 2 && t.front() == '\'' && t.back() == '\'') {
    out.clear();
    for (size_t i = 1; i + 1 < t.size(); i++) {
      out += t[i];
      if (t[i] == '\'' && i + 2 < t.size() && t[i + 1] == '\'') i++;
    }
    return true;
  }
  if (!t.empty() && (t.front() == '"' || t.front() == '\'')) {
    err = "unterminated quoted scalar";
    return false;
  }
  if (!t.empty() && (t.front() == '&' || t.front() == '*' || t.front() == '!' || t.front() == '{' ||
                     t.front() == '|' || t.front() == '>')) {
    err = "unsupported YAML construct '" + t.substr(0, 1) + "'";
    return false;
  }
  out = t;
  return true;
}

std::shared_ptr<Node> scalarNode(const std::string& token, std::string& err) {
  auto n = std::make_shared<Node>();
  const std::string t = trim(token);
  if (t.empty() || t == "~" || t == "null" || t == "Null" || t == "NULL") return n;  // null
  if (t.front() == '[') {
    if (t.back() != ']') {
      err = "unterminated flow sequence";
      return nullptr;
    }
    n->kind = Node::kSeq;
    const std::string body = t.substr(1, t.size() - 2);
    std::string cur;
    char q = 0;
    auto push = [&]() -> bool {
      const std::string item = trim(cur);
      cur.clear();
      if (item.empty()) return true;
      if (item.front() == '[') {
        err = "nested flow sequences are not supported";
        return false;
      }
      auto c = scalarNode(item, err);
      if (!c) return false;
      n->seq.push_back(c);
      return true;
    };
    for (char c : body) {
      if (q) {
        if (c == q) q = 0;
        cur += c;
      } else if (c == '"' || c == '\'') {
        q = c;
        cur += c;
      } else if (c == ',') {
        if (!push()) return nullptr;
      } else {
        cur += c;
      }
    }
    if (!push()) return nullptr;
    return n;
  }
  n->kind = Node::kScalar;
  if (!unquote(t, n->scalar, err)) return nullptr;
  return n;
}

// Split "key: value" (the first ": " or a trailing ':' outside quotes).
bool splitKey(const std::string& s, std::string& key, std::string& rest) {
  char q = 0;
  for (size_t i = 0; i < s.size(); i++) {
    const char c = s[i];
    if (q) {
      if (c == q) q = 0;
      continue;
    }
    if ((c == '"' || c == '\'') && i == 0) {
      q = c;
      continue;
    }
    if (c == ':' && (i + 1 == s.size() || isBlank(s[i + 1]))) {
      key = trim(s.substr(0, i));
      rest = trim(s.substr(i + 1));
      return true;
    }
  }
  return false;
}

struct Parser {
  std::vector<Line> lines;
  size_t pos = 0;
  std::string err;

  bool fail(int no, const std::string& m) {
    if (err.empty()) err = "line " + std::to_string(no) + ": " + m;
    return false;
  }

  bool isSeqItem(const Line& l) const { return l.text == "-" || (l.text.size() >= 2 && l.text[0] == '-' && l.text[1] == ' '); }

  // The value after "key:" or "- " when it is empty: a nested block, or null.
  std::shared_ptr<Node> nested(int parent_indent, bool allow_same_indent_seq) {
    if (pos < lines.size()) {
      const Line& l = lines[pos];
      if (l.indent > parent_indent || (allow_same_indent_seq && l.indent == parent_indent && isSeqItem(l)))
        return block(l.indent);
    }
    return std::make_shared<Node>();
  }

  std::shared_ptr<Node> block(int indent) {
    if (pos >= lines.size()) return std::make_shared<Node>();
    if (isSeqItem(lines[pos])) return sequence(indent);
    return mapping(indent);
  }

  std::shared_ptr<Node> sequence(int indent) {
    auto n = std::make_shared<Node>();
    n->kind = Node::kSeq;
    while (pos < lines.size() && lines[pos].indent == indent && isSeqItem(lines[pos])) {
      Line& l = lines[pos];
      const std::string content = l.text.size() > 1 ? trim(l.text.substr(2)) : std::string();
      if (content.empty()) {
        pos++;
        auto c = nested(indent, false);
        if (!c) return nullptr;
        n->seq.push_back(c);
        continue;
      }
      std::string key, rest;
      if (content.front() != '[' && splitKey(content, key, rest)) {
        // "- key: value": a mapping whose entries sit at the content column
        size_t k = 1;
        while (k < l.text.size() && l.text[k] == ' ') k++;
        const int col = indent + static_cast<int>(k);
        l.indent = col;
        l.text = content;
        auto c = mapping(col);
        if (!c) return nullptr;
        n->seq.push_back(c);
        continue;
      }
      auto c = scalarNode(content, err);
      if (!c) {
        fail(l.no, err.empty() ? "bad sequence item" : err);
        return nullptr;
      }
      n->seq.push_back(c);
      pos++;
    }
    if (pos < lines.size() && lines[pos].indent > indent) {
      fail(lines[pos].no, "unexpected indentation");
      return nullptr;
    }
    return n;
  }

  std::shared_ptr<Node> mapping(int indent) {
    auto n = std::make_shared<Node>();
    n->kind = Node::kMap;
    while (pos < lines.size() && lines[pos].indent == indent) {
      const Line& l = lines[pos];
      if (isSeqItem(l)) {
        fail(l.no, "sequence item inside a mapping");
        return nullptr;
      }
      std::string key, rest;
      if (!splitKey(l.text, key, rest)) {
        fail(l.no, "expected 'key: value'");
        return nullptr;
      }
      std::string k;
      std::string e;
      if (!unquote(key, k, e)) {
        fail(l.no, e);
        return nullptr;
      }
      for (const auto& kv : n->map)
        if (kv.first == k) {
          fail(l.no, "duplicate key '" + k + "'");
          return nullptr;
        }
      pos++;
      std::shared_ptr<Node> v;
      if (rest.empty()) {
        v = nested(indent, true);
      } else {
        v = scalarNode(rest, err);
        if (!v) {
          fail(l.no, err.empty() ? "bad value" : err);
          return nullptr;
        }
      }
      if (!v) return nullptr;
      n->map.emplace_back(k, v);
    }
    if (pos < lines.size() && lines[pos].indent > indent) {
      fail(lines[pos].no, "unexpected indentation");
      return nullptr;
    }
    return n;
  }
};

}  // namespace

const Node* Node::get(const std::string& key) const {
  if (kind != kMap) return nullptr;
  for (const auto& kv : map)
    if (kv.first == key) return kv.second.get();
  return nullptr;
}

std::shared_ptr<Node> parse(const std::string& text, std::string& err) {
  Parser p;
  size_t start = 0;
  int no = 0;
  while (start <= text.size()) {
    size_t end = text.find('\n', start);
    if (end == std::string::npos) end = text.size();
    std::string raw = text.substr(start, end - start);
    no++;
    start = end + 1;
    if (!raw.empty() && raw.back() == '\r') raw.pop_back();
    if (raw == "---" || raw == "...") continue;
    const std::string s = stripComment(raw);
    int indent = 0;
    while (indent < static_cast<int>(s.size()) && s[indent] == ' ') indent++;
    if (indent < static_cast<int>(s.size()) && s[indent] == '\t') {
      err = "line " + std::to_string(no) + ": tab in indentation";
      return nullptr;
    }
    const std::string t = trim(s);
    if (t.empty()) continue;
    p.lines.push_back(Line{no, indent, t});
    if (end == text.size()) break;
  }
  if (p.lines.empty()) return std::make_shared<Node>();
  auto root = p.block(p.lines[0].indent);
  if (!root || p.pos != p.lines.size()) {
    if (p.err.empty()) p.fail(p.pos < p.lines.size() ? p.lines[p.pos].no : no, "unexpected content");
    err = p.err;
    return nullptr;
  }
  return root;
}

bool asBool(const Node* n, bool fallback) {
  if (!n || !n->isScalar()) return fallback;
  const std::string& s = n->scalar;
  if (s == "true" || s == "True" || s == "TRUE" || s == "yes" || s == "Yes" || s == "YES" || s == "on" ||
      s == "On" || s == "ON")
    return true;
  if (s == "false" || s == "False" || s == "FALSE" || s == "no" || s == "No" || s == "NO" || s == "off" ||
      s == "Off" || s == "OFF")
    return false;
  return fallback;
}

int64_t asInt(const Node* n, int64_t fallback) {
  if (!n || !n->isScalar() || n->scalar.empty()) return fallback;
  errno = 0;
  char* end = nullptr;
  const long long v = std::strtoll(n->scalar.c_str(), &end, 0);
  if (errno || *end) return fallback;
  return v;
}

uint64_t asUInt(const Node* n, uint64_t fallback) {
  if (!n || !n->isScalar() || n->scalar.empty() || n->scalar[0] == '-') return fallback;
  errno = 0;
  char* end = nullptr;
  const unsigned long long v = std::strtoull(n->scalar.c_str(), &end, 0);
  if (errno || *end) return fallback;
  return v;
}

double asDouble(const Node* n, double fallback) {
  if (!n || !n->isScalar() || n->scalar.empty()) return fallback;
  const std::string& s = n->scalar;
  if (s == ".inf" || s == ".Inf" || s == ".INF" || s == "+.inf") return HUGE_VAL;
  if (s == "-.inf" || s == "-.Inf" || s == "-.INF") return -HUGE_VAL;
  if (s == ".nan" || s == ".NaN" || s == ".NAN") return std::nan("");
  errno = 0;
  char* end = nullptr;
  const double v = std::strtod(s.c_str(), &end);
  if (errno == ERANGE || *end) return fallback;
  return v;
}

std::string asString(const Node* n, const std::string& fallback) {
  if (!n || !n->isScalar()) return fallback;
  return n->scalar;
}

}  // namespace yaml
}  // namespace misa
