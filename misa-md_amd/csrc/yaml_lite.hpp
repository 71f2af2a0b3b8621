// A reader for the subset of YAML that MISA-MD configuration files use:
// block mappings and block sequences by indentation, flow sequences
// ("[1, 2, 3]"), plain and quoted scalars, and "#" comments.  It stands in
// for yaml-cpp's YAML::Load in the frontend's ConfigParser
// (frontend/config_parser.cpp:46-82); anchors, tags, multi-line scalars and
// flow mappings are not supported and are reported as errors.
#pragma once
#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <vector>

namespace misa {
namespace yaml {

struct Node {
  enum Kind { kNull, kScalar, kMap, kSeq } kind = kNull;
  std::string scalar;
  std::vector<std::pair<std::string, std::shared_ptr<Node>>> map;  // in file order
  std::vector<std::shared_ptr<Node>> seq;

  bool isMap() const { return kind == kMap; }
  bool isSeq() const { return kind == kSeq; }
  bool isScalar() const { return kind == kScalar; }
  // child of a mapping, or null when absent (as yaml-cpp's operator[])
  const Node* get(const std::string& key) const;
  size_t size() const { return kind == kSeq ? seq.size() : (kind == kMap ? map.size() : 0); }
  const Node* at(size_t i) const { return i < seq.size() ? seq[i].get() : nullptr; }
};

// Parse `text`; on failure returns null and sets `err` (with a line number).
std::shared_ptr<Node> parse(const std::string& text, std::string& err);

// Scalar conversions in the manner of yaml-cpp's as<T>(fallback): the
// fallback when the node is absent or does not convert.
bool asBool(const Node* n, bool fallback);
int64_t asInt(const Node* n, int64_t fallback);
uint64_t asUInt(const Node* n, uint64_t fallback);
double asDouble(const Node* n, double fallback);
std::string asString(const Node* n, const std::string& fallback);

}  // namespace yaml
}  // namespace misa
