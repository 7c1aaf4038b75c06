// yaml_reader.cpp — block-style YAML subset reader for material files.
#include "yaml_reader.h"

#include <cstdlib>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace uepm::pseudopotential::yaml {

namespace {

struct Line {
    int indent;
    std::string text;
    int number;
};

std::string trim(const std::string& s) {
    size_t a = s.find_first_not_of(" \t\r");
    if (a == std::string::npos) return "";
    size_t b = s.find_last_not_of(" \t\r");
    return s.substr(a, b - a + 1);
}

std::string unquote(std::string v) {
    if (v.size() >= 2 && ((v.front() == '"' && v.back() == '"') || (v.front() == '\'' && v.back() == '\'')))
        return v.substr(1, v.size() - 2);
    return v;
}

std::string strip_comment(const std::string& s) {
    bool sq = false, dq = false;
    for (size_t i = 0; i < s.size(); ++i) {
        const char c = s[i];
        if (c == '\'' && !dq) sq = !sq;
        if (c == '"' && !sq) dq = !dq;
        if (c == '#' && !sq && !dq && (i == 0 || s[i - 1] == ' ' || s[i - 1] == '\t')) return s.substr(0, i);
    }
    return s;
}

[[noreturn]] void error(int line, const std::string& what) {
    throw std::runtime_error("YAML parse error at line " + std::to_string(line) + ": " + what);
}

class Parser {
 public:
    explicit Parser(std::vector<Line> lines) : L(std::move(lines)) {}

    std::shared_ptr<Node> block(size_t& i, int indent) {
        auto n = std::make_shared<Node>();
        if (i >= L.size()) return n;
        if (is_item(L[i])) return seq(i, indent);
        n->kind = Node::Map;
        while (i < L.size() && L[i].indent == indent && !is_item(L[i])) {
            const Line& ln = L[i];
            const size_t c = key_colon(ln.text);
            if (c == std::string::npos) error(ln.number, "expected 'key: value'");
            const std::string key = unquote(trim(ln.text.substr(0, c)));
            const std::string val = trim(ln.text.substr(c + 1));
            ++i;
            std::shared_ptr<Node> child;
            if (!val.empty()) {
                if (val.front() == '{' || val.front() == '[') error(ln.number, "flow collections are not supported");
                child = std::make_shared<Node>();
                child->kind = Node::Scalar;
                child->scalar = unquote(val);
            } else if (i < L.size() && L[i].indent > indent) {
                child = block(i, L[i].indent);
            } else if (i < L.size() && L[i].indent == indent && is_item(L[i])) {
                child = seq(i, indent);
            } else {
                child = std::make_shared<Node>();
            }
            n->map.emplace_back(key, child);
        }
        if (i < L.size() && L[i].indent > indent) error(L[i].number, "unexpected indentation");
        return n;
    }

 private:
    std::vector<Line> L;

    static bool is_item(const Line& l) { return l.text == "-" || l.text.rfind("- ", 0) == 0; }

    static size_t key_colon(const std::string& t) {
        for (size_t i = 0; i < t.size(); ++i)
            if (t[i] == ':' && (i + 1 == t.size() || t[i + 1] == ' ')) return i;
        return std::string::npos;
    }

    std::shared_ptr<Node> seq(size_t& i, int indent) {
        auto n = std::make_shared<Node>();
        n->kind = Node::Seq;
        while (i < L.size() && L[i].indent == indent && is_item(L[i])) {
            Line& ln = L[i];
            const std::string rest = trim(ln.text.substr(1));
            if (rest.empty()) {
                ++i;
                if (i < L.size() && L[i].indent > indent)
                    n->seq.push_back(block(i, L[i].indent));
                else
                    n->seq.push_back(std::make_shared<Node>());
                continue;
            }
            if (key_colon(rest) == std::string::npos) {  // scalar item
                auto s = std::make_shared<Node>();
                s->kind = Node::Scalar;
                s->scalar = unquote(rest);
                n->seq.push_back(s);
                ++i;
                continue;
            }
            // "- key: value": the mapping continues at the column of `key`
            const int inner = indent + static_cast<int>(ln.text.size() - rest.size());
            ln.indent = inner;
            ln.text = rest;
            n->seq.push_back(block(i, inner));
        }
        return n;
    }
};

}  // namespace

const Node* Node::get(const std::string& key) const {
    if (kind != Map) return nullptr;
    for (const auto& kv : map)
        if (kv.first == key) return kv.second.get();
    return nullptr;
}

double Node::as_double(const std::string& where) const {
    if (kind != Scalar) throw std::runtime_error("YAML: " + where + " is not a scalar");
    char* end = nullptr;
    const double v = std::strtod(scalar.c_str(), &end);
    if (end == scalar.c_str() || trim(end) != "") throw std::runtime_error("YAML: " + where + " is not a number: " + scalar);
    return v;
}

std::shared_ptr<Node> parse(const std::string& text) {
    std::vector<Line> lines;
    std::istringstream in(text);
    std::string raw;
    int number = 0;
    while (std::getline(in, raw)) {
        ++number;
        if (raw.find('\t') != std::string::npos && trim(raw).size() && raw.find_first_not_of(" \t") > 0 &&
            raw.substr(0, raw.find_first_not_of(" \t")).find('\t') != std::string::npos)
            error(number, "tab indentation");
        const std::string body = strip_comment(raw);
        const std::string t = trim(body);
        if (t.empty() || t == "---" || t == "...") continue;
        const int indent = static_cast<int>(body.find_first_not_of(' '));
        lines.push_back({indent, t, number});
    }
    if (lines.empty()) return std::make_shared<Node>();
    Parser p(lines);
    size_t i = 0;
    auto root = p.block(i, lines[0].indent);
    if (i != lines.size()) error(lines[i].number, "unexpected content");
    return root;
}

std::shared_ptr<Node> load_file(const std::string& filename) {
    std::ifstream f(filename);
    if (!f) throw std::runtime_error("Materials::load_material_parameters: cannot open " + filename);
    std::stringstream ss;
    ss << f.rdbuf();
    return parse(ss.str());
}

}  // namespace uepm::pseudopotential::yaml
