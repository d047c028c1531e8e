// A small block-YAML reader for AMCA3D input decks.
//
// The reference parses its input with yaml-cpp 0.3 (Cafe.cpp:111-121) and reads
// values through get_if_present / get_if_present_no_default / get_required and
// expect_* (CafeParsing.h:97-160, CafeParsing.cpp:86-145).  This reader covers
// the subset the decks use -- block maps and sequences, "- key: value" items,
// flow sequences [a, b], quoted and plain scalars, '#' comments -- and keeps
// yaml-cpp 0.3's scalar conversions: numbers through a stream with base
// detection for integers, booleans from y/n, yes/no, true/false, on/off in
// lower, Capitalised or UPPER case.
#pragma once
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace amca_host {
namespace yaml {

enum class NodeType { Null = 0, Scalar = 1, Sequence = 2, Map = 3 };

class Node {
 public:
  NodeType type = NodeType::Null;
  std::string scalar;
  std::vector<Node> seq;
  std::vector<std::pair<std::string, Node>> map;
  int line = 0;  // 1-based source line, for messages

  NodeType Type() const { return type; }
  size_t size() const {
    return type == NodeType::Sequence ? seq.size() : type == NodeType::Map ? map.size() : 0;
  }
  const Node& operator[](size_t i) const;
  const Node* FindValue(const std::string& key) const;
  std::string describe() const;
};

Node parse(const std::string& text);
Node parse_file(const std::string& path);

// scalar conversions (throw std::runtime_error on failure)
void convert(const Node& n, std::string& out);
void convert(const Node& n, double& out);
void convert(const Node& n, int& out);
void convert(const Node& n, long long& out);
void convert(const Node& n, bool& out);
void convert(const Node& n, unsigned& out);
// three-component lists: operator>>(node, double*) of CafeParsing.cpp:66-82
// reads elements 0, 1 and, when present, 2
void convert3(const Node& n, double* out);
void convert3(const Node& n, int* out);

template <class T>
void get_if_present(const Node& node, const std::string& key, T& result,
                    const T& default_if_not_present = T()) {
  const Node* v = node.FindValue(key);
  if (v) convert(*v, result);
  else result = default_if_not_present;
}

template <class T>
void get_if_present_no_default(const Node& node, const std::string& key, T& result) {
  const Node* v = node.FindValue(key);
  if (v) convert(*v, result);
}

template <class T>
void get_required(const Node& node, const std::string& key, T& result) {
  const Node* v = node.FindValue(key);
  if (!v) throw std::runtime_error("Error: parsing missing required key: " + key);
  convert(*v, result);
}

void get_required3(const Node& node, const std::string& key, double* out);
void get_required3(const Node& node, const std::string& key, int* out);
void get_if_present_no_default3(const Node& node, const std::string& key, double* out);

const Node* expect_type(const Node& node, const std::string& key, NodeType type, bool optional);
inline const Node* expect_sequence(const Node& n, const std::string& k, bool optional = false) {
  return expect_type(n, k, NodeType::Sequence, optional);
}
inline const Node* expect_map(const Node& n, const std::string& k, bool optional = false) {
  return expect_type(n, k, NodeType::Map, optional);
}
inline const Node* expect_scalar(const Node& n, const std::string& k, bool optional = false) {
  return expect_type(n, k, NodeType::Scalar, optional);
}

}  // namespace yaml
}  // namespace amca_host
