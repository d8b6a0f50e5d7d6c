// Minimal XML reader for the parameter files: elements, attributes,
// self-closing tags, comments, declarations and processing instructions.
// Text content is ignored (the parameter schema keeps every value in an
// attribute).  Missing attributes read as 0 / false, as the reference's
// XML accessors return (toolkit.h:137-271).
#pragma once

#include <memory>
#include <string>
#include <utility>
#include <vector>

namespace hamx {
namespace xml {

struct Element {
  std::string name;
  std::vector<std::pair<std::string, std::string>> attributes;
  std::vector<std::unique_ptr<Element>> children;

  // first child with the given name, or nullptr
  const Element *child(const std::string &n) const;
  // every child with the given name, in document order
  std::vector<const Element *> children_named(const std::string &n) const;
  // attribute text, or nullptr when absent
  const std::string *attribute(const std::string &n) const;
  double attr_double(const std::string &n) const;
  unsigned long long attr_uint(const std::string &n) const;
  bool attr_bool(const std::string &n) const;
  std::string attr_string(const std::string &n) const; // throws when absent
};

struct Document {
  std::unique_ptr<Element> root;
};

// Parse text / a file; throws std::runtime_error on malformed input.
Document parse(const std::string &text);
Document load(const std::string &filename);

// Follow a chain of first-child names from `e`; throws std::runtime_error
// naming the missing element.
const Element *trace(const Element *e, const std::vector<std::string> &chain);

} // namespace xml
} // namespace hamx
