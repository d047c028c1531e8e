#include "yaml_lite.h"

#include <cctype>
#include <cstdlib>
#include <fstream>

namespace amca_host {
namespace yaml {

namespace {

struct Line {
  int indent;
  std::string text;
  int lineno;
};

std::string trim(const std::string& s) {
  size_t a = 0, b = s.size();
  while (a < b && std::isspace((unsigned char)s[a])) a++;
  while (b > a && std::isspace((unsigned char)s[b - 1])) b--;
  return s.substr(a, b - a);
}

// Removes a trailing comment: '#' at the start or after whitespace, outside quotes.
std::string strip_comment(const std::string& s) {
  char q = 0;
  for (size_t i = 0; i < s.size(); i++) {
    char c = s[i];
    if (q) {
      if (c == q) q = 0;
      continue;
    }
    if (c == '\'' || c == '"') q = c;
    else if (c == '#' && (i == 0 || std::isspace((unsigned char)s[i - 1]))) return s.substr(0, i);
  }
  return s;
}

// Position of the ':' that separates a map key from its value, or npos.
size_t key_colon(const std::string& s) {
  char q = 0;
  int depth = 0;
  for (size_t i = 0; i < s.size(); i++) {
    char c = s[i];
    if (q) {
      if (c == q) q = 0;
      continue;
    }
    if (c == '\'' || c == '"') q = c;
    else if (c == '[' || c == '{') depth++;
    else if (c == ']' || c == '}') depth--;
    else if (c == ':' && depth == 0 && (i + 1 == s.size() || s[i + 1] == ' ' || s[i + 1] == '\t'))
      return i;
  }
  return std::string::npos;
}

bool is_seq_item(const std::string& t) { return t == "-" || (t.size() >= 2 && t[0] == '-' && t[1] == ' '); }

std::string unquote(const std::string& s) {
  if (s.size() >= 2 && ((s.front() == '"' && s.back() == '"') || (s.front() == '\'' && s.back() == '\'')))
    return s.substr(1, s.size() - 2);
  return s;
}

Node scalar_node(const std::string& text, int lineno) {
  Node n;
  n.line = lineno;
  std::string t = trim(text);
  if (t.empty() || t == "~" || t == "null" || t == "Null" || t == "NULL") {
    n.type = NodeType::Null;
    return n;
  }
  n.type = NodeType::Scalar;
  n.scalar = unquote(t);
  return n;
}

// flow collections: [a, b, [c]] and {k: v}
Node parse_flow(const std::string& s, size_t& i, int lineno);

void skip_ws(const std::string& s, size_t& i) {
  while (i < s.size() && std::isspace((unsigned char)s[i])) i++;
}

Node parse_flow_scalar(const std::string& s, size_t& i, int lineno) {
  skip_ws(s, i);
  size_t a = i;
  char q = 0;
  while (i < s.size()) {
    char c = s[i];
    if (q) {
      if (c == q) q = 0;
      i++;
      continue;
    }
    if (c == '\'' || c == '"') {
      q = c;
      i++;
      continue;
    }
    if (c == ',' || c == ']' || c == '}') break;
    i++;
  }
  return scalar_node(s.substr(a, i - a), lineno);
}

Node parse_flow(const std::string& s, size_t& i, int lineno) {
  skip_ws(s, i);
  Node n;
  n.line = lineno;
  if (i < s.size() && s[i] == '[') {
    n.type = NodeType::Sequence;
    i++;
    skip_ws(s, i);
    if (i < s.size() && s[i] == ']') {
      i++;
      return n;
    }
    while (i < s.size()) {
      skip_ws(s, i);
      Node item = (s[i] == '[' || s[i] == '{') ? parse_flow(s, i, lineno) : parse_flow_scalar(s, i, lineno);
      n.seq.push_back(item);
      skip_ws(s, i);
      if (i < s.size() && s[i] == ',') {
        i++;
        continue;
      }
      if (i < s.size() && s[i] == ']') {
        i++;
        return n;
      }
      break;
    }
    throw std::runtime_error("yaml: unterminated flow sequence at line " + std::to_string(lineno));
  }
  if (i < s.size() && s[i] == '{') {
    n.type = NodeType::Map;
    i++;
    skip_ws(s, i);
    if (i < s.size() && s[i] == '}') {
      i++;
      return n;
    }
    while (i < s.size()) {
      skip_ws(s, i);
      size_t a = i;
      while (i < s.size() && s[i] != ':' && s[i] != ',' && s[i] != '}') i++;
      std::string key = unquote(trim(s.substr(a, i - a)));
      Node val;
      val.line = lineno;
      if (i < s.size() && s[i] == ':') {
        i++;
        skip_ws(s, i);
        val = (s[i] == '[' || s[i] == '{') ? parse_flow(s, i, lineno) : parse_flow_scalar(s, i, lineno);
      }
      n.map.emplace_back(key, val);
      skip_ws(s, i);
      if (i < s.size() && s[i] == ',') {
        i++;
        continue;
      }
      if (i < s.size() && s[i] == '}') {
        i++;
        return n;
      }
      break;
    }
    throw std::runtime_error("yaml: unterminated flow map at line " + std::to_string(lineno));
  }
  return parse_flow_scalar(s, i, lineno);
}

Node inline_value(const std::string& text, int lineno) {
  std::string t = trim(text);
  if (!t.empty() && (t[0] == '[' || t[0] == '{')) {
    size_t i = 0;
    Node n = parse_flow(t, i, lineno);
    skip_ws(t, i);
    if (i != t.size())
      throw std::runtime_error("yaml: trailing text after flow collection at line " + std::to_string(lineno));
    return n;
  }
  return scalar_node(t, lineno);
}

class Parser {
 public:
  explicit Parser(std::vector<Line> lines) : L(std::move(lines)) {}

  Node document() {
    if (L.empty()) return Node();
    size_t i = 0;
    Node n = block(i, L[0].indent);
    if (i != L.size())
      throw std::runtime_error("yaml: unexpected indentation at line " + std::to_string(L[i].lineno));
    return n;
  }

 private:
  std::vector<Line> L;

  Node block(size_t& i, int indent) {
    const Line& ln = L[i];
    if (is_seq_item(ln.text)) return sequence(i, indent);
    if (key_colon(ln.text) != std::string::npos) return mapping(i, indent);
    Node n = inline_value(ln.text, ln.lineno);
    i++;
    return n;
  }

  Node mapping(size_t& i, int indent) {
    Node n;
    n.type = NodeType::Map;
    n.line = L[i].lineno;
    while (i < L.size() && L[i].indent == indent && !is_seq_item(L[i].text)) {
      const Line ln = L[i];
      size_t c = key_colon(ln.text);
      if (c == std::string::npos)
        throw std::runtime_error("yaml: expected 'key: value' at line " + std::to_string(ln.lineno));
      std::string key = unquote(trim(ln.text.substr(0, c)));
      std::string rest = trim(ln.text.substr(c + 1));
      i++;
      Node val;
      val.line = ln.lineno;
      if (!rest.empty()) {
        val = inline_value(rest, ln.lineno);
      } else if (i < L.size() && L[i].indent > indent) {
        val = block(i, L[i].indent);
      } else if (i < L.size() && L[i].indent == indent && is_seq_item(L[i].text)) {
        val = sequence(i, indent);
      }
      n.map.emplace_back(key, std::move(val));
    }
    return n;
  }

  Node sequence(size_t& i, int indent) {
    Node n;
    n.type = NodeType::Sequence;
    n.line = L[i].lineno;
    while (i < L.size() && L[i].indent == indent && is_seq_item(L[i].text)) {
      const Line ln = L[i];
      std::string after = ln.text.size() > 1 ? ln.text.substr(1) : std::string();
      size_t lead = 0;
      while (lead < after.size() && after[lead] == ' ') lead++;
      std::string rest = after.substr(lead);
      if (rest.empty()) {
        i++;
        if (i < L.size() && L[i].indent > indent) n.seq.push_back(block(i, L[i].indent));
        else n.seq.push_back(Node());
        continue;
      }
      const int item_indent = indent + 1 + (int)lead;
      if (is_seq_item(rest) || key_colon(rest) != std::string::npos) {
        // the item is a collection whose first line starts after "- "
        L[i].indent = item_indent;
        L[i].text = rest;
        n.seq.push_back(block(i, item_indent));
      } else {
        n.seq.push_back(inline_value(rest, ln.lineno));
        i++;
      }
    }
    return n;
  }
};

bool flexible_case(const std::string& s) {
  if (s.empty()) return true;
  bool all_lower = true, all_upper = true, first_cap = std::isupper((unsigned char)s[0]) != 0;
  for (size_t i = 0; i < s.size(); i++) {
    if (std::isupper((unsigned char)s[i])) all_lower = false;
    if (std::islower((unsigned char)s[i])) all_upper = false;
    if (i > 0 && std::isupper((unsigned char)s[i])) first_cap = false;
  }
  return all_lower || all_upper || first_cap;
}

const Node& scalar_of(const Node& n, const char* what) {
  if (n.type != NodeType::Scalar)
    throw std::runtime_error(std::string("yaml: expected a scalar for ") + what + " at line " +
                             std::to_string(n.line));
  return n;
}

template <class T>
void convert_number(const Node& n, T& out, const char* what) {
  const Node& s = scalar_of(n, what);
  std::stringstream ss(s.scalar);
  ss.unsetf(std::ios::dec);
  T v;
  if (!(ss >> v) || !(ss >> std::ws).eof())
    throw std::runtime_error("yaml: invalid scalar '" + s.scalar + "' for " + what + " at line " +
                             std::to_string(n.line));
  out = v;
}

}  // namespace

const Node& Node::operator[](size_t i) const {
  if (type != NodeType::Sequence || i >= seq.size())
    throw std::runtime_error("yaml: bad sequence index at line " + std::to_string(line));
  return seq[i];
}

const Node* Node::FindValue(const std::string& key) const {
  if (type != NodeType::Map) return nullptr;
  for (const auto& kv : map)
    if (kv.first == key) return &kv.second;
  return nullptr;
}

std::string Node::describe() const {
  static const char* names[] = {"Null", "Scalar", "Sequence", "Map"};
  return std::string(names[(int)type]) + " at line " + std::to_string(line);
}

Node parse(const std::string& text) {
  std::vector<Line> lines;
  std::istringstream in(text);
  std::string raw;
  int no = 0;
  while (std::getline(in, raw)) {
    no++;
    if (!raw.empty() && raw.back() == '\r') raw.pop_back();
    if (raw == "---" || raw == "...") continue;
    std::string s = strip_comment(raw);
    for (char& c : s)
      if (c == '\t') c = ' ';
    std::string t = trim(s);
    if (t.empty()) continue;
    int ind = 0;
    while (ind < (int)s.size() && s[ind] == ' ') ind++;
    lines.push_back({ind, t, no});
  }
  return Parser(std::move(lines)).document();
}

Node parse_file(const std::string& path) {
  std::ifstream f(path);
  if (!f.good()) throw std::runtime_error("Input file is not specified or does not exist: " + path);
  std::stringstream ss;
  ss << f.rdbuf();
  return parse(ss.str());
}

void convert(const Node& n, std::string& out) { out = scalar_of(n, "string").scalar; }
void convert(const Node& n, double& out) { convert_number(n, out, "double"); }
void convert(const Node& n, int& out) { convert_number(n, out, "int"); }
void convert(const Node& n, long long& out) { convert_number(n, out, "long"); }
void convert(const Node& n, unsigned& out) { convert_number(n, out, "unsigned"); }

void convert(const Node& n, bool& out) {
  const Node& s = scalar_of(n, "bool");
  static const char* names[4][2] = {{"y", "n"}, {"yes", "no"}, {"true", "false"}, {"on", "off"}};
  if (flexible_case(s.scalar)) {
    std::string lower = s.scalar;
    for (char& c : lower) c = (char)std::tolower((unsigned char)c);
    for (auto& p : names) {
      if (lower == p[0]) {
        out = true;
        return;
      }
      if (lower == p[1]) {
        out = false;
        return;
      }
    }
  }
  throw std::runtime_error("yaml: invalid boolean '" + s.scalar + "' at line " + std::to_string(n.line));
}

void convert3(const Node& n, double* out) {
  convert(n[0], out[0]);
  convert(n[1], out[1]);
  if (n.size() > 2) convert(n[2], out[2]);
}

void convert3(const Node& n, int* out) {
  convert(n[0], out[0]);
  convert(n[1], out[1]);
  if (n.size() > 2) convert(n[2], out[2]);
}

void get_required3(const Node& node, const std::string& key, double* out) {
  const Node* v = node.FindValue(key);
  if (!v) throw std::runtime_error("Error: parsing missing required key: " + key);
  convert3(*v, out);
}

void get_required3(const Node& node, const std::string& key, int* out) {
  const Node* v = node.FindValue(key);
  if (!v) throw std::runtime_error("Error: parsing missing required key: " + key);
  convert3(*v, out);
}

void get_if_present_no_default3(const Node& node, const std::string& key, double* out) {
  const Node* v = node.FindValue(key);
  if (v) convert3(*v, out);
}

const Node* expect_type(const Node& node, const std::string& key, NodeType type, bool optional) {
  const Node* v = node.FindValue(key);
  if (!optional && !v)
    throw std::runtime_error("Error: parsing expected required value " + key + " but it was not found");
  if (v && v->type != type)
    throw std::runtime_error("Error: parsing - Check indentation of input file. (key " + key + ": " +
                             v->describe() + ")");
  return v;
}

}  // namespace yaml
}  // namespace amca_host
