// yaml_reader.h — a reader for the block-style YAML subset of the material
// parameter files: nested mappings by indentation, sequences of mappings
// ("- key: value"), scalars, comments and a "---" document marker.
#pragma once

#include <map>
#include <memory>
#include <string>
#include <vector>

namespace uepm::pseudopotential::yaml {

struct Node {
    enum Kind { Null, Scalar, Map, Seq } kind = Null;
    std::string scalar;
    std::vector<std::pair<std::string, std::shared_ptr<Node>>> map;
    std::vector<std::shared_ptr<Node>> seq;

    const Node* get(const std::string& key) const;  // nullptr when absent
    double as_double(const std::string& where) const;
    std::string as_string() const { return scalar; }
};

/* Parses `text`; throws std::runtime_error with a line number on malformed input. */
std::shared_ptr<Node> parse(const std::string& text);
std::shared_ptr<Node> load_file(const std::string& filename);

}  // namespace uepm::pseudopotential::yaml
